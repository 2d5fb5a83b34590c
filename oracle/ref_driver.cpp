// Golden-vector driver for the real OpenPhase library (test infrastructure).
//
// Built only by oracle/Makefile.ref against the reference sources under
// /root/reference (compiled from source into oracle/_ref/, never copied).
// It runs the canonical grain-growth loop of examples/NormalGG/NormalGG.cpp:74-85
//   IP.Set -> DO.CalculatePhaseFieldIncrements -> Phi.NormalizeIncrements -> Phi.MergeIncrements
// (optionally with the driving-force calls of benchmarks/AnisotropyTest/AnisotropyTest.cpp:140-155)
// on a microstructure read from a raw PhaseField .dat file (src/PhaseField.cpp:2340-2388 format)
// and writes the finalized state at selected steps, in the same raw format but without the
// zero-valued entries that CalculateDerivativesSR appends (src/PhaseField.cpp:566-615).
//
// usage: ref_driver <input.opi> <state.dat> <grains.bin> <nsteps> <dt> <outprefix> <dumpsteps,...>
//        [--driving-force] [--threads N] [--write-files] [--normalize-only <pairs.bin>]
//        [--energy] [--dg-every N]
//        [--untimed N]
// --untimed N: the first N steps run but are left out of the per-call timers (warm-up).
// --normalize-only: after FinalizeInitialization, adds the pair increments of pairs.bin to
// Phi.FieldsDot with NodeAB::add_asym1 (uint32 ncells; per cell uint32 linear index (x outer,
// z inner), uint32 npairs, npairs x {uint32 a; uint32 b; double value}), runs only
// Phi.NormalizeIncrements(BC, dt) and writes <outprefix>_pairs.bin: per listed cell uint32 npairs,
// then npairs x {uint32 indexA; uint32 indexB; double value1} as FieldsDot holds them.
// --write-files: at the dump steps also the reference's own writers: PhaseField::Write (full lists,
// zero-valued entries included) to <outprefix>_<step>_full.dat, GrainsProperties::Write to
// <outprefix>_<step>.grains and PhaseField::WriteVTK into the settings' VTK directory.
// grains.bin: uint32 n; n x { uint32 phase; double q[4] }   (q = s, x, y, z)
// Writes <outprefix>_<step>.dat and <outprefix>_<step>.vol (n doubles Volume, n doubles VolumeRatio)
// and prints one JSON line with per-call wall times, InterfaceProperties::ReportMaximumTimeStep
// before the first IP.Set ("max_dt_init") and after the IP.Set of every step ("max_dt", one per
// step). The limiter's plausibility warnings go to stderr as the reference writes them.
// --energy: also DoubleObstacle::AverageEnergyDensity after every step ("energy") and after one
// more IP.Set on the final state ("energy_ip_current").
// --dg-every N: with --driving-force, UserDrivingForce::SetDrivingForce only in steps divisible
// by N (DrivingForce::Clear, Average and MergePhaseFieldIncrements in every step).
// --regularization: the InterfaceRegularization overload of the step
// (src/DoubleObstacle.cpp:644-684, 686-770): DO.CalculatePhaseFieldIncrements(Phi, IP[, dG], Kappa),
// then Kappa.Average(Phi, BC) and Kappa.MergePhaseFieldIncrements(Phi, IP)
// (src/InterfaceRegularization.cpp:156-176, 403-456), before the driving-force calls.

#include "Settings.h"
#include "RunTimeControl.h"
#include "DoubleObstacle.h"
#include "PhaseField.h"
#include "BoundaryConditions.h"
#include "InterfaceProperties.h"
#include "DrivingForce.h"
#include "UserDrivingForce.h"
#include "InterfaceRegularization.h"

#include <cstdio>
#include <cstdint>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>
#include <omp.h>

using namespace openphase;

static void dump_state(const PhaseField& Phi, const std::string& prefix, int step)
{
    std::string fn = prefix + "_" + std::to_string(step) + ".dat";
    std::ofstream out(fn, std::ios::binary);
    int nx = Phi.Grid.Nx, ny = Phi.Grid.Ny, nz = Phi.Grid.Nz;
    out.write(reinterpret_cast<const char*>(&nx), sizeof(int));
    out.write(reinterpret_cast<const char*>(&ny), sizeof(int));
    out.write(reinterpret_cast<const char*>(&nz), sizeof(int));
    for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
    for (long k = 0; k < nz; ++k)
    {
        const NodePF& node = Phi.Fields(i, j, k);
        size_t n = 0;
        for (auto it = node.cbegin(); it != node.cend(); ++it) if (it->value != 0.0) ++n;
        out.write(reinterpret_cast<const char*>(&n), sizeof(size_t));
        for (auto it = node.cbegin(); it != node.cend(); ++it)
        if (it->value != 0.0)
        {
            size_t idx = it->index;
            double v = it->value;
            out.write(reinterpret_cast<const char*>(&idx), sizeof(size_t));
            out.write(reinterpret_cast<const char*>(&v), sizeof(double));
        }
    }
    std::string fv = prefix + "_" + std::to_string(step) + ".vol";
    std::ofstream ov(fv, std::ios::binary);
    size_t ng = Phi.FieldsProperties.size();
    for (size_t g = 0; g < ng; ++g) ov.write(reinterpret_cast<const char*>(&Phi.FieldsProperties[g].Volume), sizeof(double));
    for (size_t g = 0; g < ng; ++g) ov.write(reinterpret_cast<const char*>(&Phi.FieldsProperties[g].VolumeRatio), sizeof(double));
}

int main(int argc, char* argv[])
{
    if (argc < 8)
    {
        std::fprintf(stderr, "usage: %s input.opi state.dat grains.bin nsteps dt outprefix dumpsteps [--driving-force] [--threads N]\n", argv[0]);
        return 2;
    }
    std::string opi = argv[1], state = argv[2], grains = argv[3];
    int nsteps = std::atoi(argv[4]);
    double dt = std::strtod(argv[5], nullptr);
    std::string prefix = argv[6];
    std::set<int> dumps;
    {
        std::stringstream ss(argv[7]);
        std::string tok;
        while (std::getline(ss, tok, ',')) if (!tok.empty()) dumps.insert(std::atoi(tok.c_str()));
    }
    bool with_dg = false, write_files = false, energy = false, with_reg = false;
    std::string normalize_only;
    int threads = 1, untimed = 0, dg_every = 1;
    for (int a = 8; a < argc; ++a)
    {
        std::string s = argv[a];
        if (s == "--driving-force") with_dg = true;
        if (s == "--write-files") write_files = true;
        if (s == "--threads" && a + 1 < argc) threads = std::atoi(argv[++a]);
        if (s == "--normalize-only" && a + 1 < argc) normalize_only = argv[++a];
        if (s == "--untimed" && a + 1 < argc) untimed = std::atoi(argv[++a]);
        if (s == "--energy") energy = true;
        if (s == "--regularization") with_reg = true;
        if (s == "--dg-every" && a + 1 < argc) dg_every = std::atoi(argv[++a]);
    }
    omp_set_num_threads(threads);

    Settings OPSettings;
    OPSettings.ReadInput(opi);
    PhaseField Phi(OPSettings, opi);
    DoubleObstacle DO(OPSettings, opi);
    InterfaceProperties IP(OPSettings, opi);
    BoundaryConditions BC(OPSettings, opi);
    DrivingForce* dG = nullptr;
    UserDrivingForce* UDF = nullptr;
    if (with_dg)
    {
        dG = new DrivingForce(OPSettings, opi);
        UDF = new UserDrivingForce(OPSettings, opi);
    }
    InterfaceRegularization* Kappa = with_reg ? new InterfaceRegularization(OPSettings, opi) : nullptr;

    std::ifstream gin(grains, std::ios::binary);
    uint32_t ng = 0;
    gin.read(reinterpret_cast<char*>(&ng), sizeof(uint32_t));
    for (uint32_t g = 0; g < ng; ++g)
    {
        uint32_t phase = 0;
        double q[4];
        gin.read(reinterpret_cast<char*>(&phase), sizeof(uint32_t));
        gin.read(reinterpret_cast<char*>(q), 4 * sizeof(double));
        size_t idx = Phi.AddGrainInfo(phase);
        Phi.FieldsProperties[idx].Orientation.set(q[0], q[1], q[2], q[3]);
    }
    if (!Phi.Read(state)) return 3;
    Phi.FinalizeInitialization(BC);
    if (dumps.count(0)) dump_state(Phi, prefix, 0);

    if (!normalize_only.empty())
    {
        std::ifstream pin(normalize_only, std::ios::binary);
        uint32_t nc = 0;
        pin.read(reinterpret_cast<char*>(&nc), sizeof(uint32_t));
        const long NY = Phi.Grid.Ny, NZ = Phi.Grid.Nz;
        std::vector<uint32_t> cells(nc);
        for (uint32_t c = 0; c < nc; ++c)
        {
            uint32_t lin = 0, np = 0;
            pin.read(reinterpret_cast<char*>(&lin), sizeof(uint32_t));
            pin.read(reinterpret_cast<char*>(&np), sizeof(uint32_t));
            cells[c] = lin;
            const long i = lin / (NY * NZ), j = (lin / NZ) % NY, k = lin % NZ;
            for (uint32_t q = 0; q < np; ++q)
            {
                uint32_t a = 0, b = 0;
                double v = 0.0;
                pin.read(reinterpret_cast<char*>(&a), sizeof(uint32_t));
                pin.read(reinterpret_cast<char*>(&b), sizeof(uint32_t));
                pin.read(reinterpret_cast<char*>(&v), sizeof(double));
                Phi.FieldsDot(i, j, k).add_asym1(a, b, v);
            }
        }
        Phi.NormalizeIncrements(BC, dt);
        std::ofstream pout(prefix + "_pairs.bin", std::ios::binary);
        for (uint32_t lin : cells)
        {
            const long i = lin / (NY * NZ), j = (lin / NZ) % NY, k = lin % NZ;
            const auto& node = Phi.FieldsDot(i, j, k);
            uint32_t np = static_cast<uint32_t>(node.size());
            pout.write(reinterpret_cast<const char*>(&np), sizeof(uint32_t));
            for (auto it = node.cbegin(); it != node.cend(); ++it)
            {
                uint32_t a = static_cast<uint32_t>(it->indexA), b = static_cast<uint32_t>(it->indexB);
                double v = it->value1;
                pout.write(reinterpret_cast<const char*>(&a), sizeof(uint32_t));
                pout.write(reinterpret_cast<const char*>(&b), sizeof(uint32_t));
                pout.write(reinterpret_cast<const char*>(&v), sizeof(double));
            }
        }
        std::printf("{\"cells\": %ld, \"normalized\": %u}\n", (long)Phi.Grid.Nx * Phi.Grid.Ny * Phi.Grid.Nz, nc);
        return 0;
    }

    const double max_dt_init = IP.ReportMaximumTimeStep();
    std::vector<double> max_dt, energies;
    double t_ip = 0, t_do = 0, t_norm = 0, t_merge = 0, t_dg = 0;
    for (int step = 1; step <= nsteps; ++step)
    {
        double t0 = omp_get_wtime();
        if (with_dg) dG->Clear();
        IP.Set(Phi, BC);
        double t1 = omp_get_wtime();
        if (with_reg)
        {
            if (with_dg) DO.CalculatePhaseFieldIncrements(Phi, IP, *dG, *Kappa);
            else         DO.CalculatePhaseFieldIncrements(Phi, IP, *Kappa);
            Kappa->Average(Phi, BC);
            Kappa->MergePhaseFieldIncrements(Phi, IP);
        }
        else if (with_dg) DO.CalculatePhaseFieldIncrements(Phi, IP, *dG);
        else              DO.CalculatePhaseFieldIncrements(Phi, IP);
        double t2 = omp_get_wtime();
        if (with_dg)
        {
            if (step % dg_every == 0) UDF->SetDrivingForce(Phi, *dG);
            dG->Average(Phi, BC);
            dG->MergePhaseFieldIncrements(Phi, IP);
        }
        double t3 = omp_get_wtime();
        Phi.NormalizeIncrements(BC, dt);
        double t4 = omp_get_wtime();
        Phi.MergeIncrements(BC, dt);
        double t5 = omp_get_wtime();
        if (step > untimed) { t_ip += t1 - t0; t_do += t2 - t1; t_dg += t3 - t2; t_norm += t4 - t3; t_merge += t5 - t4; }
        max_dt.push_back(IP.ReportMaximumTimeStep());
        if (energy) energies.push_back(DO.AverageEnergyDensity(Phi, IP));
        if (dumps.count(step)) dump_state(Phi, prefix, step);
        if (dumps.count(step) && write_files)
        {
            Phi.Write(prefix + "_" + std::to_string(step) + "_full.dat");
            Phi.FieldsProperties.Write(prefix + "_" + std::to_string(step) + ".grains");
            Phi.WriteVTK(OPSettings, step);
        }
    }
    long cells = (long)Phi.Grid.Nx * Phi.Grid.Ny * Phi.Grid.Nz;
    double total = t_ip + t_do + t_dg + t_norm + t_merge;
    std::printf("{\"cells\": %ld, \"steps\": %d, \"threads\": %d, \"IP.Set\": %.6f, \"CalculatePhaseFieldIncrements\": %.6f, "
                "\"DrivingForce\": %.6f, \"NormalizeIncrements\": %.6f, \"MergeIncrements\": %.6f, \"total_s\": %.6f, "
                "\"cell_updates_per_s\": %.6e, \"max_dt_init\": %.17g, \"max_dt\": [",
                cells, nsteps - untimed, threads, t_ip, t_do, t_dg, t_norm, t_merge, total,
                total > 0 ? double(cells) * (nsteps - untimed) / total : 0.0, max_dt_init);
    for (size_t i = 0; i < max_dt.size(); ++i) std::printf("%s%.17g", i ? ", " : "", max_dt[i]);
    std::printf("]");
    if (energy)
    {
        // DoubleObstacle::AverageEnergyDensity after every step (IP from the step's IP.Set), then
        // once more after a fresh IP.Set on the final state
        std::printf(", \"energy\": [");
        for (size_t i = 0; i < energies.size(); ++i) std::printf("%s%.17g", i ? ", " : "", energies[i]);
        IP.Set(Phi, BC);
        std::printf("], \"energy_ip_current\": %.17g", DO.AverageEnergyDensity(Phi, IP));
    }
    std::printf("}\n");
    delete dG;
    delete UDF;
    return 0;
}
