// Test driver of include/openphase_amd/io.h: turns a finalized state (own lists, flags, grain
// records) into the reference's three output files, whose bytes the tests compare with the
// reference's own output for the same state.
//
// usage: io_driver Nx Ny Nz nphases state.dat flags.bin grains.bin outprefix
//   Nx Ny Nz    input sizes (0 = reduced axis)
//   state.dat   own (non-zero) lists in the raw phase-field format
//   flags.bin   int8 flag per cell
//   grains.bin  uint32 n, then n x {uint8 exist; int32 stage, phase, state; double refvolume,
//               volume, maxvolume, q[4]} (packed)
// writes outprefix_full.dat, outprefix.grains, outprefix.vts
#include "openphase_amd/io.h"

#include <cstdlib>
#include <iostream>

using namespace openphase_amd::io;

int main(int argc, char** argv)
{
    if (argc != 9) {
        std::cerr << "usage: io_driver Nx Ny Nz nphases state.dat flags.bin grains.bin outprefix\n";
        return 2;
    }
    const Dims d = Dims::from_input(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]));
    const int nphases = std::atoi(argv[4]);
    Lists own;
    if (!read_phasefield(argv[5], d, own)) { std::cerr << "state\n"; return 3; }
    std::vector<int8_t> flag(static_cast<size_t>(d.cells()));
    {
        std::ifstream f(argv[6], std::ios::binary);
        f.read(reinterpret_cast<char*>(flag.data()), static_cast<std::streamsize>(flag.size()));
        if (!f) { std::cerr << "flags\n"; return 3; }
    }
    std::vector<GrainRecord> grains;
    {
        std::ifstream f(argv[7], std::ios::binary);
        uint32_t n = 0;
        f.read(reinterpret_cast<char*>(&n), sizeof(n));
        grains.resize(n);
        for (GrainRecord& g : grains) {
            uint8_t exist = 0;
            int32_t stage = 0, phase = 0, state = 0;
            f.read(reinterpret_cast<char*>(&exist), 1);
            f.read(reinterpret_cast<char*>(&stage), 4);
            f.read(reinterpret_cast<char*>(&phase), 4);
            f.read(reinterpret_cast<char*>(&state), 4);
            f.read(reinterpret_cast<char*>(&g.ref_volume), 8);
            f.read(reinterpret_cast<char*>(&g.volume), 8);
            f.read(reinterpret_cast<char*>(&g.max_volume), 8);
            f.read(reinterpret_cast<char*>(g.orientation), 32);
            g.exist = exist != 0;
            g.stage = stage;
            g.phase = static_cast<size_t>(phase);
            g.state = state;
        }
        if (!f) { std::cerr << "grains\n"; return 3; }
    }
    const std::string out = argv[8];
    const Lists full = full_lists(d, own, flag);
    if (!write_phasefield(out + "_full.dat", d, full)) return 4;
    if (!write_grains(out + ".grains", grains)) return 4;
    if (!write_vtk(out + ".vts", d, nphases, grains, full, flag)) return 4;
    return 0;
}
