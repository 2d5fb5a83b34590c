// OpenPhaseAMD.h -- OpenPhase's DoubleObstacle / PhaseField / InterfaceProperties /
// BoundaryConditions call surface on top of the MI355X step (C ABI include/mpf.h).
//
// A grain-growth driver written against the reference (examples/NormalGG/NormalGG.cpp:58-85)
//
//     Settings OPSettings;  OPSettings.ReadInput(InputFile);
//     PhaseField Phi(OPSettings, InputFile);          DoubleObstacle DO(OPSettings, InputFile);
//     InterfaceProperties IP(OPSettings, InputFile);  BoundaryConditions BC(OPSettings, InputFile);
//     ... Phi.Read(file) / grains ... ; Phi.FinalizeInitialization(BC);
//     for (...) { IP.Set(Phi, BC); DO.CalculatePhaseFieldIncrements(Phi, IP);
//                 Phi.NormalizeIncrements(BC, dt); Phi.MergeIncrements(BC, dt); }
//
// compiles unchanged against this header (namespace openphase_amd instead of openphase). The
// four calls of a step are recorded and executed as one fused device step when MergeIncrements
// is called; calling them out of order, or asking for the intermediate increments, is an error.
// Errors are reported by exceptions (std::runtime_error); nothing calls exit().
//
// Input: the same .opi files (keys of @GridParameters, @Settings, @PhaseField,
// @InterfaceProperties, @BoundaryConditions, @UserDrivingForce, @DrivingForce,
// @InterfaceRegularization read by the reference's FileInterface, src/FileInterface.cpp:129-345).
#ifndef OPENPHASE_AMD_OPENPHASEAMD_H
#define OPENPHASE_AMD_OPENPHASEAMD_H

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <initializer_list>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mpf.h"
#include "io.h"

namespace openphase_amd {

// ------------------------------------------------------------------------------------------
// .opi reader: "@Module" opens a block, "$Key  free text  : value" defines a key in it.
class InputFile {
 public:
    explicit InputFile(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("input file \"" + path + "\" could not be opened");
        std::string line, module;
        while (std::getline(in, line)) {
            size_t p = line.find_first_not_of(" \t\r");
            if (p == std::string::npos) continue;
            if (line[p] == '@') {
                std::istringstream ss(line.substr(p + 1));
                ss >> module;
                entries_[module];  // a module without keys still counts as present
                continue;
            }
            if (line[p] != '$') continue;
            std::istringstream ss(line.substr(p + 1));
            std::string key;
            ss >> key;
            size_t colon = line.find(':', p);
            if (colon == std::string::npos) continue;
            std::string value = line.substr(colon + 1);
            size_t a = value.find_first_not_of(" \t\r");
            size_t b = value.find_last_not_of(" \t\r");
            value = a == std::string::npos ? std::string() : value.substr(a, b - a + 1);
            if (!entries_[module].count(key)) entries_[module][key] = value;
        }
    }
    bool has_module(const std::string& module) const { return entries_.count(module) != 0; }
    bool has(const std::string& module, const std::string& key) const
    {
        auto m = entries_.find(module);
        return m != entries_.end() && m->second.count(key);
    }
    std::string get(const std::string& module, const std::string& key, const std::string& def = std::string(),
                    bool mandatory = false) const
    {
        if (has(module, key)) return entries_.at(module).at(key);
        if (mandatory) throw std::runtime_error("missing $" + key + " in @" + module);
        return def;
    }
    double getd(const std::string& module, const std::string& key, double def, bool mandatory = false) const
    {
        std::string v = get(module, key, std::string(), mandatory);
        return v.empty() ? def : std::stod(v);
    }
    int geti(const std::string& module, const std::string& key, int def, bool mandatory = false) const
    {
        std::string v = get(module, key, std::string(), mandatory);
        return v.empty() ? def : std::stoi(v);
    }
    bool getb(const std::string& module, const std::string& key, bool def) const
    {
        std::string v = get(module, key);
        if (v.empty()) return def;
        for (auto& ch : v) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
        return v == "YES" || v == "TRUE" || v == "1";
    }

 private:
    std::map<std::string, std::map<std::string, std::string>> entries_;
};

inline void check(mpf_status s, mpf_ctx* ctx, const char* where)
{
    if (s == MPF_OK) return;
    std::string msg = std::string(where) + " failed (status " + std::to_string(static_cast<int>(s)) + ")";
    if (ctx) msg += ": " + std::string(mpf_error_message(ctx));
    throw std::runtime_error(msg);
}

// ------------------------------------------------------------------------------------------
// GridParameters (include/GridParameters.h:37-172) and Settings (include/Settings.h:44-119)
struct GridParameters {
    int Nx = 0, Ny = 0, Nz = 0;   // as given in the input (0 = reduced axis)
    double dx = 1.0;
    double iWidth = 5.0;
    double Eta() const { return iWidth * dx; }
    int Active() const { return (Nx != 0) + (Ny != 0) + (Nz != 0); }
    long TotalNumberOfCells() const
    {
        return static_cast<long>(Nx ? Nx : 1) * (Ny ? Ny : 1) * (Nz ? Nz : 1);
    }
    // volume in 3D, area in 2D, length in 1D of one cell
    double CellVolume(bool ignore_active_dimensions = false) const
    {
        return std::pow(dx, ignore_active_dimensions ? 3 : Active());
    }
};

class Settings {
 public:
    void ReadInput(const std::string& file)
    {
        InputFile in(file);
        Grid.Nx = in.geti("GridParameters", "Nx", 0, true);
        Grid.Ny = in.geti("GridParameters", "Ny", 0, true);
        Grid.Nz = in.geti("GridParameters", "Nz", 0, true);
        Grid.dx = in.getd("GridParameters", "dx", 0.0, true);
        Grid.iWidth = in.has("GridParameters", "IWidth") ? in.getd("GridParameters", "IWidth", 5.0)
                                                          : in.getd("GridParameters", "InterfaceWidth", 5.0, true);
        PhaseNames.clear();
        for (size_t n = 0; in.has("Settings", "Phase_" + std::to_string(n)); ++n)
            PhaseNames.push_back(in.get("Settings", "Phase_" + std::to_string(n)));
        Nphases = PhaseNames.size();
        Solid.assign(Nphases, 1);
        for (size_t n = 0; n < Nphases; ++n)
            Solid[n] = in.get("Settings", "State_" + std::to_string(n), "SOLID") == "SOLID" ? 1 : 0;
        // Settings::ReadInput (src/Settings.cpp:245-247): output directories
        VTKDir = in.get("Settings", "VTKDir", "VTK/");
        RawDataDir = in.get("Settings", "RAWDir", "RawData/");
        InputFileName = file;
    }
    GridParameters Grid;
    std::string VTKDir = "VTK/";
    std::string RawDataDir = "RawData/";
    size_t Nphases = 0;
    std::vector<std::string> PhaseNames;
    std::vector<uint8_t> Solid;
    std::string InputFileName;
};

// ------------------------------------------------------------------------------------------
// RunTimeControl (include/RunTimeControl.h; keys of RunTimeControl::ReadInput,
// src/RunTimeControl.cpp:99-131): the time loop's step counter, number of steps and time step.
class RunTimeControl {
 public:
    RunTimeControl() = default;
    RunTimeControl(Settings& s, const std::string& file) { (void)s; ReadInput(file); }
    void ReadInput(const std::string& file)
    {
        InputFile in(file);
        const char* m = "RunTimeControl";
        nSteps = in.has(m, "MaxTimeStep") ? in.geti(m, "MaxTimeStep", 0) : in.geti(m, "nSteps", 0, true);
        dt = in.getd(m, "dt", 0.0, true);
        tStart = in.has(m, "StartTimeStep") ? in.geti(m, "StartTimeStep", 0) : in.geti(m, "tStart", 0);
        tStep = tStart;
        FTime = in.has(m, "VTKOutputInterval") ? in.geti(m, "VTKOutputInterval", 1) : in.geti(m, "FTime", 1);
        STime = in.has(m, "ConsoleOutputInterval") ? in.geti(m, "ConsoleOutputInterval", 1) : in.geti(m, "STime", 1);
        RTime = in.has(m, "CheckpointInterval") ? in.geti(m, "CheckpointInterval", 1) : in.geti(m, "tRstrt", 1);
    }
    void IncrementTimeStep() { ++tStep; }
    bool WriteVTK() const { return FTime > 0 && tStep % FTime == 0; }
    bool WriteToScreen() const { return STime > 0 && tStep % STime == 0; }
    bool WriteRawData() const { return RTime > 0 && tStep % RTime == 0; }
    int tStep = 0, tStart = 0, nSteps = 0;
    int FTime = 1, STime = 1, RTime = 1;
    double dt = 0.0;
};

// ------------------------------------------------------------------------------------------
class BoundaryConditions {
 public:
    BoundaryConditions() = default;
    BoundaryConditions(Settings& s, const std::string& file) { ReadInput(file); (void)s; }
    void ReadInput(const std::string& file)
    {
        InputFile in(file);
        const char* keys[6] = {"BC0X", "BCNX", "BC0Y", "BCNY", "BC0Z", "BCNZ"};
        for (int i = 0; i < 6; ++i) {
            std::string v = in.get("BoundaryConditions", keys[i], "Periodic");
            if (v == "Periodic") bc[i] = MPF_BC_PERIODIC;
            else if (v == "NoFlux") bc[i] = MPF_BC_NOFLUX;
            else if (v == "Free") bc[i] = MPF_BC_FREE;
            else if (v == "Fixed") bc[i] = MPF_BC_FIXED;
            else if (v == "Mirror") bc[i] = MPF_BC_MIRROR;
            else throw std::runtime_error("unknown boundary condition " + v);
        }
    }
    int32_t bc[6] = {0, 0, 0, 0, 0, 0};
};

class InterfaceProperties;
class DoubleObstacle;
class DrivingForce;
class UserDrivingForce;

// ------------------------------------------------------------------------------------------
// PhaseField (include/PhaseField.h): owns the device state through an mpf_ctx.
class PhaseField {
 public:
    PhaseField(Settings& s, const std::string& file, int precision = MPF_FP64, int slots = 8)
        : Grid(s.Grid), Nphases(s.Nphases), settings_(s), file_(file), precision_(precision), slots_(slots)
    {
        InputFile in(file);
        ConsiderNucleusVolume = in.getb("PhaseField", "ConsiderNucleusVolume", true);
        NucleusVolumeFactor = in.getd("PhaseField", "NucleusVolumeFactor", 1.0);
    }
    ~PhaseField()
    {
        if (ctx_) mpf_destroy(ctx_);
    }
    PhaseField(const PhaseField&) = delete;
    PhaseField& operator=(const PhaseField&) = delete;

    // PhaseField::AddGrainInfo (src/PhaseField.cpp:2970-2979); `quaternion` (s, x, y, z) is the
    // Orientation the grain table file stores (rotation9 is what the step uses)
    size_t AddGrainInfo(size_t PhaseIndex, const double* rotation9 = nullptr, const double* quaternion = nullptr)
    {
        grain_phase_.push_back(static_cast<int32_t>(PhaseIndex));
        for (int q = 0; q < 9; ++q) grain_rot_.push_back(rotation9 ? rotation9[q] : (q % 4 == 0 ? 1.0 : 0.0));
        for (int q = 0; q < 4; ++q) grain_quat_.push_back(quaternion ? quaternion[q] : 0.0);
        grains_dirty_ = true;
        return grain_phase_.size() - 1;
    }
    size_t NumberOfGrains() const { return grain_phase_.size(); }

    // PhaseField::Read(FileName) (src/PhaseField.cpp:2487-2543): raw .dat
    bool Read(const std::string& FileName)
    {
        std::ifstream inp(FileName, std::ios::binary);
        if (!inp) return false;
        int n[3];
        inp.read(reinterpret_cast<char*>(n), sizeof(n));
        const int sx = Grid.Nx ? Grid.Nx : 1, sy = Grid.Ny ? Grid.Ny : 1, sz = Grid.Nz ? Grid.Nz : 1;
        if (n[0] != sx || n[1] != sy || n[2] != sz) return false;
        const long N = static_cast<long>(sx) * sy * sz;
        std::vector<int32_t> count(N), index;
        std::vector<double> value;
        for (long c = 0; c < N; ++c) {
            uint64_t k = 0;
            inp.read(reinterpret_cast<char*>(&k), sizeof(k));
            count[c] = static_cast<int32_t>(k);
            for (uint64_t q = 0; q < k; ++q) {
                uint64_t idx = 0;
                double v = 0.0;
                inp.read(reinterpret_cast<char*>(&idx), sizeof(idx));
                inp.read(reinterpret_cast<char*>(&v), sizeof(v));
                index.push_back(static_cast<int32_t>(idx));
                value.push_back(v);
            }
        }
        if (!inp) return false;
        Upload(count.data(), index.data(), value.data());
        return true;
    }
    // Lists in the raw order (count per cell, concatenated index/value)
    void Upload(const int32_t* count, const int32_t* index, const double* value)
    {
        ensure_context();
        check(mpf_upload(ctx_, count, index, value), ctx_, "PhaseField::Read");
        fractions_step_ = -1;
    }
    // PhaseField::Write(FileName) (src/PhaseField.cpp:2340-2388): the lists as the reference holds
    // them after a step, zero-valued appended entries included
    bool Write(const std::string& FileName) const
    {
        if (!ctx_) return false;
        io::Lists own;
        std::vector<int8_t> flag;
        Download(own.count, own.index, own.value, &flag);
        return io::write_phasefield(FileName, dims(), io::full_lists(dims(), own, flag));
    }
    // PhaseField::Write(Settings, tStep) (:2390-2401): PhaseField_<step>.dat and
    // GrainsProperties_<step>.dat in RawDataDir
    bool Write(const Settings& s, const int tStep) const
    {
        return Write(io::make_file_name(s.RawDataDir, "PhaseField_", tStep, ".dat")) &&
               FieldsProperties.Write(io::make_file_name(s.RawDataDir, "GrainsProperties_", tStep, ".dat"));
    }
    // PhaseField::WriteVTK (src/PhaseField.cpp:2128-2158), single resolution: VTKDir/PhaseField_<step>.vts
    void WriteVTK(Settings& s, const int tStep, const bool CurvatureOutput = false, const int precision = 16) const
    {
        if (CurvatureOutput) throw std::runtime_error("PhaseField::WriteVTK: curvature output is not implemented");
        io::Lists own;
        std::vector<int8_t> flag;
        Download(own.count, own.index, own.value, &flag);
        const std::vector<double>& frac = Fractions.values();
        const std::string fn = io::make_file_name(s.VTKDir, "PhaseField_", tStep, ".vts");
        if (!io::write_vtk(fn, dims(), static_cast<int>(Nphases), grain_records(), io::full_lists(dims(), own, flag),
                           flag, precision, &frac))
            throw std::runtime_error("PhaseField::WriteVTK: cannot write " + fn);
    }
    // GrainsProperties of the phase field: Write(FileName) (include/GrainsProperties.h:768-795)
    struct GrainsTable {
        const PhaseField* owner;
        bool Write(const std::string& FileName) const { return io::write_grains(FileName, owner->grain_records()); }
        size_t size() const { return owner->grain_phase_.size(); }
    };
    GrainsTable FieldsProperties{this};
    // PhaseField::Fractions (src/PhaseField.cpp:411-423, kept current by every Finalize): phase
    // fractions per cell, computed on the device from the current state when first read after a
    // step; Fractions(i, j, k, {n})
    struct FractionsField {
        const PhaseField* owner;
        double operator()(int i, int j, int k, std::initializer_list<size_t> n) const
        {
            const GridParameters& G = owner->Grid;
            const long ny = G.Ny ? G.Ny : 1, nz = G.Nz ? G.Nz : 1;
            const long cell = (static_cast<long>(i) * ny + j) * nz + k;
            return values()[*n.begin() * static_cast<size_t>(G.TotalNumberOfCells()) + static_cast<size_t>(cell)];
        }
        // [phase * cells + cell]
        const std::vector<double>& values() const
        {
            PhaseField* p = const_cast<PhaseField*>(owner);
            if (p->fractions_step_ != p->steps_ || p->fractions_.empty()) {
                p->fractions_.assign(p->Nphases * static_cast<size_t>(p->Grid.TotalNumberOfCells()), 0.0);
                check(mpf_phase_fractions(p->ctx_, p->fractions_.data(), 1), p->ctx_, "PhaseField::Fractions");
                p->fractions_step_ = p->steps_;
            }
            return p->fractions_;
        }
    };
    FractionsField Fractions{this};
    // Grain records as GrainsProperties holds them (device volumes, stages; host phase, orientation)
    std::vector<io::GrainRecord> grain_records() const
    {
        const size_t G = grain_phase_.size();
        std::vector<io::GrainRecord> g(G);
        if (!ctx_ || G == 0) return g;
        std::vector<double> vol(G), mv(G), rv(G);
        std::vector<uint8_t> ex(G);
        std::vector<int32_t> st(G);
        check(mpf_get_grains(ctx_, vol.data(), nullptr), ctx_, "GrainsProperties");
        check(mpf_get_grain_records(ctx_, mv.data(), rv.data(), ex.data(), st.data()), ctx_, "GrainsProperties");
        for (size_t i = 0; i < G; ++i) {
            g[i].exist = ex[i] != 0;
            g[i].stage = st[i];
            g[i].phase = static_cast<size_t>(grain_phase_[i]);
            g[i].state = static_cast<size_t>(grain_phase_[i]) < settings_.Solid.size() && !settings_.Solid[grain_phase_[i]] ? 1 : 0;
            g[i].ref_volume = rv[i];
            g[i].volume = vol[i];
            g[i].max_volume = mv[i];
            for (int q = 0; q < 4; ++q) g[i].orientation[q] = grain_quat_[4 * i + q];
        }
        return g;
    }
    void Download(std::vector<int32_t>& count, std::vector<int32_t>& index, std::vector<double>& value,
                  std::vector<int8_t>* flag = nullptr) const
    {
        int64_t total = 0;
        check(mpf_total_entries(ctx_, &total), ctx_, "PhaseField::Download");
        count.resize(static_cast<size_t>(Grid.TotalNumberOfCells()));
        index.resize(static_cast<size_t>(total));
        value.resize(static_cast<size_t>(total));
        if (flag) flag->resize(count.size());
        check(mpf_download(ctx_, count.data(), index.data(), value.data(), flag ? flag->data() : nullptr), ctx_,
              "PhaseField::Download");
    }
    io::Dims dims() const { return io::Dims::from_input(Grid.Nx, Grid.Ny, Grid.Nz); }

    // PhaseField::FinalizeInitialization (src/PhaseField.cpp:1068-1077)
    void FinalizeInitialization(const BoundaryConditions& BC)
    {
        check_bc(BC);
        ensure_context();
        check(mpf_finalize_initialization(ctx_), ctx_, "PhaseField::FinalizeInitialization");
        fractions_step_ = -1;
        stage_ = 0;
    }
    // PhaseField::NormalizeIncrements (src/PhaseField.cpp:1334-1349): third call of a step
    void NormalizeIncrements(const BoundaryConditions& BC, const double dt)
    {
        check_bc(BC);
        if (stage_ != 2) throw std::logic_error("NormalizeIncrements before IP.Set and CalculatePhaseFieldIncrements");
        dt_ = dt;
        stage_ = 3;
    }
    // PhaseField::MergeIncrements (src/PhaseField.cpp:1248-1266) incl. Finalize: runs the step
    void MergeIncrements(const BoundaryConditions& BC, const double dt, const bool finalize = true,
                         const bool clear = true)
    {
        check_bc(BC);
        if (!finalize || !clear) throw std::logic_error("MergeIncrements without finalize/clear is not supported");
        if (stage_ != 3 || dt != dt_) throw std::logic_error("MergeIncrements must follow NormalizeIncrements with the same dt");
        // the constant driving force acts in this step only if the driver set, averaged and
        // merged it (DrivingForce::MergePhaseFieldIncrements after UserDrivingForce::SetDrivingForce
        // and DrivingForce::Average)
        if (driving_force_)
            check(mpf_set_step_terms(ctx_, dg_step_ != 0, dg_step_ == 2), ctx_, "PhaseField::MergeIncrements");
        // the averaged curvature is merged on the device: the driver must have asked for both
        // (without Average the reference merges it with weight 0, i.e. not at all)
        if (regularization_ && reg_average_ && reg_calls_ != 7)
            throw std::logic_error("a regularized step needs InterfaceRegularization::Average and "
                                   "MergePhaseFieldIncrements after CalculatePhaseFieldIncrements(Phase, IP, Kappa)");
        check(mpf_step(ctx_, dt, 1), ctx_, "PhaseField::MergeIncrements");
        stage_ = 0;
        dg_step_ = 0;
        reg_calls_ = 0;
        ++steps_;
    }
    // Runs n steps of the recorded sequence at once (one synchronisation): the calls of the last
    // step, its driving-force terms included, IP.Set in each
    void Advance(double dt, int nsteps)
    {
        check(mpf_step(ctx_, dt, nsteps), ctx_, "PhaseField::Advance");
        if (nsteps > 0) {
            ip_step_ = steps_ + nsteps - 1;
            steps_ += nsteps;
        }
    }
    // GrainsProperties Volume / VolumeRatio
    void GrainVolumes(std::vector<double>& volume, std::vector<double>& ratio) const
    {
        volume.resize(grain_phase_.size());
        ratio.resize(grain_phase_.size());
        check(mpf_get_grains(ctx_, volume.data(), ratio.data()), ctx_, "PhaseField::GrainVolumes");
    }

    GridParameters Grid;
    size_t Nphases;
    bool ConsiderNucleusVolume = true;
    double NucleusVolumeFactor = 1.0;

 private:
    friend class InterfaceProperties;
    friend class DoubleObstacle;
    friend class DrivingForce;
    friend class UserDrivingForce;
    friend class InterfaceRegularization;

    void check_bc(const BoundaryConditions& BC)
    {
        for (int i = 0; i < 6; ++i) bc_[i] = BC.bc[i];
    }
    void ensure_context();

    Settings& settings_;
    std::string file_;
    int precision_;
    int slots_;
    int32_t bc_[6] = {0, 0, 0, 0, 0, 0};
    std::vector<int32_t> grain_phase_;
    std::vector<double> grain_rot_;
    std::vector<double> grain_quat_;
    bool grains_dirty_ = true;
    mpf_ctx* ctx_ = nullptr;
    int stage_ = 0;      // 0 idle, 1 IP.Set done, 2 increments done, 3 normalized
    double dt_ = 0.0;
    bool driving_force_ = false;  // the input file gives a constant driving force (device tables)
    int dg_step_ = 0;    // driving force merged into this step: 0 none, 1 raw (SkipAverage), 2 averaged
    bool regularization_ = false;  // the input file has @InterfaceRegularization: every step takes its overload
    bool reg_average_ = true;      // its Average setting
    int reg_calls_ = 0;  // this step: 1 regularized increments, |2 Kappa.Average, |4 Kappa.MergePhaseFieldIncrements
    long steps_ = 0;     // steps run on the device
    long ip_step_ = -1;  // steps_ at the last IP.Set on this phase field (-1: none)
    std::vector<double> fractions_;  // Fractions downloaded at step fractions_step_
    long fractions_step_ = -1;
};

// ------------------------------------------------------------------------------------------
// InterfaceProperties (include/InterfaceProperties.h): phase-pair energy/mobility models.
class InterfaceProperties {
 public:
    InterfaceProperties(Settings& s, const std::string& file) : Nphases(s.Nphases), dx_(s.Grid.dx), active_(s.Grid.Active())
    {
        InputFile in(file);
        const size_t np2 = Nphases * Nphases;
        Sigma.assign(np2, 0.0);
        Mu.assign(np2, 0.0);
        EnergyModel.assign(np2, MPF_ENERGY_ISO);
        Epsilon.assign(np2, 0.0);
        for (size_t a = 0; a < Nphases; ++a)
            for (size_t b = a; b < Nphases; ++b) {
                const std::string c = "_" + std::to_string(a) + "_" + std::to_string(b);
                const std::string em = in.get("InterfaceProperties", "EnergyModel" + c, "ISO");
                const std::string mm = in.get("InterfaceProperties", "MobilityModel" + c, "ISO");
                int model = MPF_ENERGY_ISO;
                if (em == "CUBIC") model = MPF_ENERGY_CUBIC;
                else if (em != "ISO") throw std::runtime_error("EnergyModel" + c + " " + em + " is not implemented");
                if (mm != "ISO") throw std::runtime_error("MobilityModel" + c + " " + mm + " is not implemented");
                const double sig = in.getd("InterfaceProperties", "Sigma" + c, 0.0, true);
                const double mu = in.getd("InterfaceProperties", "Mu" + c, 0.0, true);
                const double eps = model == MPF_ENERGY_CUBIC ? in.getd("InterfaceProperties", "EpsilonE" + c, 0.0, true) : 0.0;
                for (auto k : {a * Nphases + b, b * Nphases + a}) {
                    Sigma[k] = sig; Mu[k] = mu; EnergyModel[k] = model; Epsilon[k] = eps;
                }
            }
        const std::string tj = in.get("InterfaceProperties", "TripleJunctionModel", "MODEL");
        if (tj == "MODEL") { TripleJunctionModel = MPF_TJ_MODEL; TripleJunctionFactor = in.getd("InterfaceProperties", "TripleJunctionFactor", 0.0); }
        else if (tj == "VALUE") { TripleJunctionModel = MPF_TJ_VALUE; TripleJunctionEnergy = in.getd("InterfaceProperties", "TripleJunctionEnergy", 0.0, true); }
        else if (tj == "NONE") TripleJunctionModel = MPF_TJ_NONE;
        else throw std::runtime_error("unknown TripleJunctionModel " + tj);
        RegularizationFactor = in.getd("InterfaceProperties", "RegularizationFactor", 1.0);
        CurvatureFactor = in.getd("InterfaceProperties", "CurvatureFactor", 1.0);
        // InterfaceProperties::Extrapolate runs inside Set (src/InterfaceProperties.cpp:575,
        // 1222-1249); ZEROTHORDER does nothing there, FIRSTORDER rewrites the energies of
        // interface-halo nodes, which this build does not model
        const std::string ex = in.get("InterfaceProperties", "ExtrapolationMode", "NONE");
        if (ex != "NONE" && ex != "ZEROTHORDER")
            throw std::runtime_error("ExtrapolationMode " + ex + " is not implemented");
    }
    // InterfaceProperties::Set (src/InterfaceProperties.cpp:557-576): first call of a step
    void Set(PhaseField& Phase, const BoundaryConditions& BC)
    {
        Phase.check_bc(BC);
        Phase.ensure_context();
        Phase.stage_ = 1;
        Phase.ip_step_ = Phase.steps_;
        phase_ = &Phase;
    }
    // InterfaceProperties::ReportMaximumTimeStep (src/InterfaceProperties.cpp:523-556): from the
    // largest energy and mobility the pairs of the interface cells had in the last IP.Set
    // (SetSR, :748-756), DBL_MAX before the first. The maxima are taken when the step runs on the
    // device (in PhaseField::MergeIncrements), so between IP.Set and MergeIncrements this still
    // reports the previous step's.
    double ReportMaximumTimeStep() const
    {
        double dt = DBL_MAX;
        if (phase_ && phase_->ctx_) check(mpf_max_time_step(phase_->ctx_, &dt), phase_->ctx_, "ReportMaximumTimeStep");
        return dt;
    }
    size_t Nphases;
    std::vector<double> Sigma, Mu, Epsilon;
    std::vector<int32_t> EnergyModel;
    int TripleJunctionModel = MPF_TJ_MODEL;
    double TripleJunctionFactor = 0.0;
    double TripleJunctionEnergy = 0.0;
    double RegularizationFactor = 1.0;
    double CurvatureFactor = 1.0;

 private:
    friend class PhaseField;
    friend class DoubleObstacle;
    double dx_;
    int active_;
    const PhaseField* phase_ = nullptr;  // the phase field of the last Set
};

// ------------------------------------------------------------------------------------------
// UserDrivingForce / DrivingForce (constant phase-pair driving force, no averaging)
class UserDrivingForce {
 public:
    UserDrivingForce(Settings& s, const std::string& file) : Nphases(s.Nphases)
    {
        InputFile in(file);
        Mode.assign(Nphases * Nphases, 0);
        Value.assign(Nphases * Nphases, 0.0);
        for (size_t n = 0; n < Nphases; ++n)
            for (size_t m = n + 1; m < Nphases; ++m) {
                const std::string c = "_" + std::to_string(n) + "_" + std::to_string(m);
                const std::string mode = in.get("UserDrivingForce", "UDF_Mode" + c, "NONE");
                if (mode == "VALUE") {
                    const double v = in.getd("UserDrivingForce", "UDF_Value" + c, 0.0, true);
                    Mode[n * Nphases + m] = Mode[m * Nphases + n] = 1;
                    Value[n * Nphases + m] = v;
                    Value[m * Nphases + n] = -v;
                } else if (mode != "NONE") {
                    throw std::runtime_error("UDF_Mode" + c + " " + mode + " is not implemented");
                }
            }
    }
    // UserDrivingForce::SetDrivingForce (src/UserDrivingForce.cpp:152-175): the VALUE pairs' raw
    // driving force into dG, which DrivingForce::Average and MergePhaseFieldIncrements then apply
    void SetDrivingForce(PhaseField& Phase, DrivingForce& dG);
    size_t Nphases;
    std::vector<int32_t> Mode;
    std::vector<double> Value;
};

class DrivingForce {
 public:
    DrivingForce(Settings& s, const std::string& file) : Nphases(s.Nphases) { Read(InputFile(file), s.Grid.iWidth); }
    // DrivingForce::Initialize + ReadInput (src/DrivingForce.cpp:45-90, 126-152): Average (default
    // Yes), Range (IWidth), Threshold (1/3 for IWidth >= 5, else IWidth/15), WeightsMode, Limit
    void Read(const InputFile& in, double iwidth)
    {
        Averaging = in.getb("DrivingForce", "Average", true);
        Range = in.geti("DrivingForce", "Range", static_cast<int>(iwidth));
        PhiThreshold = in.getd("DrivingForce", "Threshold", iwidth < 5.0 ? iwidth / 15.0 : 1.0 / 3.0);
        const std::string w = in.get("DrivingForce", "WeightsMode", "PHASEFIELDS");
        WeightsMode = w == "RANGE" ? MPF_DG_WEIGHTS_RANGE : (w == "COUNTER" ? MPF_DG_WEIGHTS_COUNTER : MPF_DG_WEIGHTS_PHASEFIELDS);
        Limit = in.getd("DrivingForce", "Limit_0_0", 0.95);
    }
    // The averaging and the merge run inside the fused step of PhaseField::MergeIncrements; these
    // calls record which of them the driver asked for in this step.
    // DrivingForce::Clear (src/DrivingForce.cpp:176-183)
    void Clear() { raw_ = averaged_ = false; }
    // DrivingForce::Average (src/DrivingForce.cpp:242-256): CollectAverage / DistributeAverage,
    // or SkipAverage when Averaging is off; without it the merged force weighs 0
    void Average(const PhaseField&, const BoundaryConditions&) { averaged_ = true; }
    // DrivingForce::MergePhaseFieldIncrements (src/DrivingForce.cpp:480-495): after the
    // double-obstacle increments, before PhaseField::NormalizeIncrements
    void MergePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties&)
    {
        if (Phase.stage_ != 2)
            throw std::logic_error("DrivingForce::MergePhaseFieldIncrements must follow CalculatePhaseFieldIncrements "
                                   "and precede NormalizeIncrements");
        Phase.dg_step_ = raw_ && averaged_ ? (Averaging ? 2 : 1) : 0;
    }
    size_t Nphases;
    bool Averaging = true;
    int Range = 5;
    double PhiThreshold = 1.0 / 3.0;
    int WeightsMode = MPF_DG_WEIGHTS_PHASEFIELDS;
    double Limit = 0.95;

 private:
    friend class UserDrivingForce;
    bool raw_ = false;       // a raw driving force was set since the last Clear
    bool averaged_ = false;  // Average ran since the last Clear
};

inline void UserDrivingForce::SetDrivingForce(PhaseField& Phase, DrivingForce& dG)
{
    bool any = false;
    for (int32_t m : Mode) any = any || m == 1;
    if (!any) return;
    Phase.ensure_context();
    if (!Phase.driving_force_)
        throw std::logic_error("UserDrivingForce::SetDrivingForce: the phase field's input file gives no driving force");
    // the reference adds the raw values of a second call up (NodeDF::add_raw); that sum is not
    // modelled here
    if (dG.raw_) throw std::logic_error("UserDrivingForce::SetDrivingForce twice without DrivingForce::Clear");
    dG.raw_ = true;
}

// ------------------------------------------------------------------------------------------
// InterfaceRegularization (include/InterfaceRegularization.h): the curvature storage lives on the
// device and is never materialised on the host; these calls record which steps of the module the
// driver asked for (the fused step runs them in PhaseField::MergeIncrements).
class InterfaceRegularization {
 public:
    // InterfaceRegularization::Initialize + ReadInput (src/InterfaceRegularization.cpp:46-118):
    // Average (default Yes), Range ((IWidth + 1)/2), Threshold (DBL_EPSILON), ClearingMode
    InterfaceRegularization(Settings& s, const std::string& file) { Read(InputFile(file), s.Grid.iWidth); }
    void Read(const InputFile& in, double iwidth)
    {
        Averaging = in.getb("InterfaceRegularization", "Average", true);
        Range = in.geti("InterfaceRegularization", "Range", static_cast<int>((iwidth + 1) / 2));
        PhiThreshold = in.getd("InterfaceRegularization", "Threshold", DBL_EPSILON);
        if (in.get("InterfaceRegularization", "ClearingMode", "AUTOMATIC") != "AUTOMATIC")
            throw std::invalid_argument("InterfaceRegularization: only ClearingMode AUTOMATIC is supported");
    }
    // InterfaceRegularization::Clear: the device storage is cleared at every step
    void Clear() {}
    // InterfaceRegularization::Average (src/InterfaceRegularization.cpp:156-176)
    void Average(const PhaseField& Phase, const BoundaryConditions&)
    {
        if (Phase.stage_ != 2 || !(Phase.reg_calls_ & 1))
            throw std::logic_error("InterfaceRegularization::Average must follow "
                                   "DoubleObstacle::CalculatePhaseFieldIncrements(Phase, IP, Kappa)");
        const_cast<PhaseField&>(Phase).reg_calls_ |= 2;
    }
    // InterfaceRegularization::MergePhaseFieldIncrements (src/InterfaceRegularization.cpp:403-456)
    void MergePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties&)
    {
        if (Phase.stage_ != 2 || !(Phase.reg_calls_ & 1))
            throw std::logic_error("InterfaceRegularization::MergePhaseFieldIncrements must follow the regularized "
                                   "increments and precede NormalizeIncrements");
        Phase.reg_calls_ |= 4;
    }
    bool Averaging = true;
    int Range = 3;
    double PhiThreshold = DBL_EPSILON;
};

// ------------------------------------------------------------------------------------------
// DoubleObstacle (include/DoubleObstacle.h)
class DoubleObstacle {
 public:
    DoubleObstacle() = default;
    DoubleObstacle(Settings&, const std::string&) {}
    // DoubleObstacle::CalculatePhaseFieldIncrements (src/DoubleObstacle.cpp:461-500)
    void CalculatePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties&)
    {
        if (Phase.stage_ != 1) throw std::logic_error("CalculatePhaseFieldIncrements before IP.Set");
        if (Phase.regularization_)
            throw std::logic_error("the input file sets up InterfaceRegularization: call "
                                   "CalculatePhaseFieldIncrements(Phase, IP[, dG], Kappa)");
        Phase.stage_ = 2;
    }
    // The overload with the driving force computes the same increments: the force acts through
    // DrivingForce::MergePhaseFieldIncrements (src/DoubleObstacle.cpp:478-500, FullAnisotropy off)
    void CalculatePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties& IP, DrivingForce&)
    {
        CalculatePhaseFieldIncrements(Phase, IP);
    }
    // The curvature overloads (src/DoubleObstacle.cpp:644-684, 686-770): increments times
    // RegularizationFactor plus the curvature InterfaceRegularization averages and merges; they
    // need the @InterfaceRegularization block in the phase field's input file (the device context
    // is set up for it)
    void CalculatePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties&, InterfaceRegularization&)
    {
        if (Phase.stage_ != 1) throw std::logic_error("CalculatePhaseFieldIncrements before IP.Set");
        if (!Phase.regularization_)
            throw std::logic_error("CalculatePhaseFieldIncrements(Phase, IP, Kappa): the phase field's input file has "
                                   "no @InterfaceRegularization block");
        Phase.stage_ = 2;
        Phase.reg_calls_ = 1;
    }
    void CalculatePhaseFieldIncrements(PhaseField& Phase, InterfaceProperties& IP, DrivingForce&,
                                       InterfaceRegularization& Kappa)
    {
        CalculatePhaseFieldIncrements(Phase, IP, Kappa);
    }
    // DoubleObstacle::AverageEnergyDensity (src/DoubleObstacle.cpp:989-1005): PointEnergy summed
    // over the grid / number of cells, with the interface energies of IP's last Set on Phase
    double AverageEnergyDensity(const PhaseField& Phase, const InterfaceProperties& IP) const
    {
        return energy_sum(Phase, IP) / static_cast<double>(Phase.Grid.TotalNumberOfCells());
    }
    // DoubleObstacle::Energy (src/DoubleObstacle.cpp:953-967): the sum times the cell volume
    double Energy(const PhaseField& Phase, const InterfaceProperties& IP) const
    {
        return energy_sum(Phase, IP) * Phase.Grid.CellVolume();
    }

 private:
    static double energy_sum(const PhaseField& Phase, const InterfaceProperties& IP);
};

// Builds the device context from the input file the first time the state is needed.
inline void PhaseField::ensure_context()
{
    if (ctx_ && !grains_dirty_) return;
    if (!ctx_) {
        InterfaceProperties ip(settings_, file_);
        std::unique_ptr<UserDrivingForce> udf;
        {
            InputFile in(file_);
            bool any = false;
            for (size_t n = 0; n < Nphases; ++n)
                for (size_t m = n + 1; m < Nphases; ++m)
                    any = any || in.has("UserDrivingForce", "UDF_Mode_" + std::to_string(n) + "_" + std::to_string(m));
            if (any) udf.reset(new UserDrivingForce(settings_, file_));
        }
        mpf_grid g{};
        g.nx = Grid.Nx; g.ny = Grid.Ny; g.nz = Grid.Nz;
        g.dx = Grid.dx; g.iwidth = Grid.iWidth;
        for (int i = 0; i < 6; ++i) g.bc[i] = bc_[i];
        g.precision = precision_;
        g.slots = slots_;
        mpf_params p{};
        p.nphases = static_cast<int32_t>(Nphases);
        p.sigma = ip.Sigma.data();
        p.mu = ip.Mu.data();
        p.energy_model = ip.EnergyModel.data();
        p.epsilon = ip.Epsilon.data();
        p.tj_model = ip.TripleJunctionModel;
        p.tj_factor = ip.TripleJunctionFactor;
        p.tj_energy = ip.TripleJunctionEnergy;
        p.consider_nucleus_volume = ConsiderNucleusVolume ? 1 : 0;
        p.nucleus_volume_factor = NucleusVolumeFactor;
        p.dg_mode = udf ? udf->Mode.data() : nullptr;
        p.dg_value = udf ? udf->Value.data() : nullptr;
        p.dg_limit = 0.95;
        if (udf) {
            DrivingForce df(settings_, file_);
            p.dg_limit = df.Limit;
            p.dg_average = df.Averaging ? 1 : 0;
            p.dg_range = df.Range;
            p.dg_threshold = df.PhiThreshold;
            p.dg_weights = df.WeightsMode;
        }
        p.regularization_factor = ip.RegularizationFactor;
        if (InputFile(file_).has_module("InterfaceRegularization")) {
            InterfaceRegularization kappa(settings_, file_);
            p.regularization = 1;
            p.curvature_factor = ip.CurvatureFactor;
            p.reg_average = kappa.Averaging ? 1 : 0;
            p.reg_range = kappa.Range;
            p.reg_threshold = kappa.PhiThreshold;
            regularization_ = true;
            reg_average_ = kappa.Averaging;
        }
        check(mpf_create(&g, &p, &ctx_), nullptr, "mpf_create");
        driving_force_ = static_cast<bool>(udf);
    }
    if (grains_dirty_) {
        if (grain_phase_.empty()) throw std::logic_error("no grains: call AddGrainInfo first");
        std::vector<uint8_t> solid(grain_phase_.size(), 1);
        for (size_t g = 0; g < grain_phase_.size(); ++g)
            if (static_cast<size_t>(grain_phase_[g]) < settings_.Solid.size()) solid[g] = settings_.Solid[grain_phase_[g]];
        check(mpf_set_grains(ctx_, static_cast<int32_t>(grain_phase_.size()), grain_phase_.data(), grain_rot_.data(),
                             solid.data()),
              ctx_, "PhaseField::AddGrainInfo");
        grains_dirty_ = false;
    }
}

inline double DoubleObstacle::energy_sum(const PhaseField& Phase, const InterfaceProperties& IP)
{
    // Properties hold what IP's last Set on this phase field evaluated: nothing before one (every
    // energy 0), the current state when no step ran since, else the state the last step started
    // from; an older Set is not held by the device state
    if (IP.phase_ != &Phase || Phase.ip_step_ < 0 || !Phase.ctx_) return 0.0;
    int32_t ip = MPF_IP_NONE;
    if (Phase.ip_step_ == Phase.steps_) ip = MPF_IP_CURRENT;
    else if (Phase.ip_step_ == Phase.steps_ - 1) ip = MPF_IP_PREVIOUS;
    else throw std::logic_error("DoubleObstacle energy: IP.Set ran before the last step's start");
    double sum = 0.0;
    check(mpf_interface_energy(Phase.ctx_, ip, &sum), Phase.ctx_, "DoubleObstacle::AverageEnergyDensity");
    return sum;
}

}  // namespace openphase_amd
#endif
