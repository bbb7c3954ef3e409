#pragma once
#include "../hip_emu.h"
