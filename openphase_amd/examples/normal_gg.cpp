// Normal grain growth through the OpenPhase-style C++ interface (include/openphase_amd/OpenPhaseAMD.h):
// the time loop of the reference's examples/NormalGG/NormalGG.cpp:74-85 (RunTimeControl loop,
// IP.Set -> DO.CalculatePhaseFieldIncrements -> Phi.NormalizeIncrements -> Phi.MergeIncrements at
// RTC.dt), on a microstructure read from a raw PhaseField .dat file instead of the reference's
// Voronoi set-up, and without its console and VTK output inside the loop.
//
// usage: normal_gg <input.opi> <state.dat> <ngrains> <nsteps> <out.dat> [outdir]
// The time step is the input file's @RunTimeControl $dt; <nsteps> steps run from its tStart.
// With outdir, also the reference's output files of the last step there: PhaseField::WriteVTK
// (VTK/PhaseField_<step>.vts) and GrainsProperties::Write (GrainsProperties.dat).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "openphase_amd/OpenPhaseAMD.h"

using namespace openphase_amd;

int main(int argc, char* argv[])
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s input.opi state.dat ngrains nsteps out.dat\n", argv[0]);
        return 2;
    }
    const std::string InputFile = argv[1];
    const int ngrains = std::atoi(argv[3]);
    const int nsteps = std::atoi(argv[4]);
    try {
        Settings OPSettings;
        OPSettings.ReadInput(InputFile);
        RunTimeControl RTC(OPSettings, InputFile);
        RTC.nSteps = RTC.tStart + nsteps - 1;
        PhaseField Phi(OPSettings, InputFile);
        DoubleObstacle DO(OPSettings, InputFile);
        InterfaceProperties IP(OPSettings, InputFile);
        BoundaryConditions BC(OPSettings, InputFile);
        for (int g = 0; g < ngrains; ++g) Phi.AddGrainInfo(0);
        if (!Phi.Read(argv[2])) {
            std::fprintf(stderr, "could not read %s\n", argv[2]);
            return 3;
        }
        Phi.FinalizeInitialization(BC);
        for (RTC.tStep = RTC.tStart; RTC.tStep <= RTC.nSteps; RTC.IncrementTimeStep()) {
            IP.Set(Phi, BC);
            DO.CalculatePhaseFieldIncrements(Phi, IP);
            Phi.NormalizeIncrements(BC, RTC.dt);
            Phi.MergeIncrements(BC, RTC.dt);
        }
        const double dt = RTC.dt;
        if (!Phi.Write(argv[5])) return 4;
        if (argc > 6) {
            OPSettings.VTKDir = std::string(argv[6]) + "/VTK/";
            Phi.WriteVTK(OPSettings, nsteps);
            if (!Phi.FieldsProperties.Write(std::string(argv[6]) + "/GrainsProperties.dat")) return 4;
        }
        std::printf("normal_gg: %d steps at dt=%.17g, ReportMaximumTimeStep %.17g\n", nsteps, dt,
                    IP.ReportMaximumTimeStep());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "normal_gg: %s\n", e.what());
        return 1;
    }
    return 0;
}
