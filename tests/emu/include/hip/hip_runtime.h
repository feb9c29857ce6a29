// CPU stand-in for <hip/hip_runtime.h> used only by the test emulator build.
#pragma once
#include "../../emu_hip.h"
