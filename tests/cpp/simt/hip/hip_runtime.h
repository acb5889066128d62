// Emulator stand-in for the HIP runtime header (tests only).
#pragma once
#include "../simt_emu.h"
