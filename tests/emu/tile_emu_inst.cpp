// tile_emu_inst.cpp -- TEST INFRASTRUCTURE: one (precision, chunk height) of the host
// emulation's stage classes, chosen by -DEMU_R=double|float -DEMU_NW=4|8.
#include "tile_emu.h"

namespace emu {
template void stage_nw<EMU_R, EMU_R, EMU_NW>(Sim &, const StageIO<EMU_R> &);
}
