// On-disk formats either side of the multiphase-field step, written from host arrays of the
// finalized state (count per cell, concatenated grain index / value lists in the raw cell order,
// flags 0/1/2 per cell): the reference's raw phase-field file, grain table and VTK output.
//
//   full_lists             the lists as the reference holds them after Finalize: own entries, then
//                          the zero-valued indices CalculateDerivativesSR appends on flagged cells
//                          (src/PhaseField.cpp:566-615; Laplacian stencil order, x outer, z inner)
//   write_phasefield       PhaseField::Write (src/PhaseField.cpp:2340-2388, NodePF::write
//                          include/Containers/NodePF.h:765-775)
//   read_phasefield        PhaseField::Read (src/PhaseField.cpp:2487-2543, NodePF::read :751-763)
//   write_grains / read_grains
//                          GrainsProperties::Write / Read (include/GrainsProperties.h:196-230,
//                          768-830), Grain fields as Grain::Clear and PhaseField::AddGrainInfo
//                          leave them (src/PhaseField.cpp:2970-2979)
//   write_vtk              PhaseField::WriteVTK, single resolution, no curvature
//                          (src/PhaseField.cpp:2128-2158; VTK::Write src/VTK.cpp:75-94 and
//                          include/VTK.h:70-81, 117-218, 353-471, 662-667)
//
// Host code only (no device types); the device state reaches these through mpf_download and
// mpf_get_grains / mpf_get_grain_records.
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace openphase_amd {
namespace io {

// Grid as the reference's GridParameters hold it: Nx, Ny, Nz >= 1 (a reduced axis has size 1 and
// d = 0).
struct Dims {
    int nx = 1, ny = 1, nz = 1;
    int dnx = 0, dny = 0, dnz = 0;
    long cells() const { return static_cast<long>(nx) * ny * nz; }
    static Dims from_input(int Nx, int Ny, int Nz)
    {
        Dims d;
        d.nx = Nx ? Nx : 1; d.ny = Ny ? Ny : 1; d.nz = Nz ? Nz : 1;
        d.dnx = Nx != 0; d.dny = Ny != 0; d.dnz = Nz != 0;
        return d;
    }
};

// Lists per cell in the raw order: count[c] entries starting at offset[c] (prefix sums of count).
struct Lists {
    std::vector<int32_t> count;
    std::vector<int32_t> index;
    std::vector<double> value;
    std::vector<long> offsets() const
    {
        std::vector<long> off(count.size() + 1, 0);
        for (size_t c = 0; c < count.size(); ++c) off[c + 1] = off[c] + count[c];
        return off;
    }
};

inline int wrap_index(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Own (non-zero) entries plus, on flagged cells, the zero-valued indices the derivative pass
// appends: every 3^d neighbour (all Laplacian weights are non-zero; the gradient stencil adds no
// new neighbour) in stencil order, each neighbour's non-zero entries in its list order.
inline Lists full_lists(const Dims& d, const Lists& own, const std::vector<int8_t>& flag)
{
    const long N = d.cells();
    const std::vector<long> off = own.offsets();
    Lists out;
    out.count.resize(N);
    for (long c = 0; c < N; ++c) {
        const long first = static_cast<long>(out.index.size());
        for (long q = off[c]; q < off[c + 1]; ++q) {
            out.index.push_back(own.index[q]);
            out.value.push_back(own.value[q]);
        }
        if (flag[c] != 0) {
            const int i = static_cast<int>(c / (static_cast<long>(d.ny) * d.nz));
            const int j = static_cast<int>((c / d.nz) % d.ny), k = static_cast<int>(c % d.nz);
            for (int di = -d.dnx; di <= d.dnx; ++di)
                for (int dj = -d.dny; dj <= d.dny; ++dj)
                    for (int dk = -d.dnz; dk <= d.dnz; ++dk) {
                        const long nb = (static_cast<long>(wrap_index(i + di, d.nx)) * d.ny + wrap_index(j + dj, d.ny)) * d.nz +
                                        wrap_index(k + dk, d.nz);
                        for (long q = off[nb]; q < off[nb + 1]; ++q) {
                            if (own.value[q] == 0.0) continue;
                            bool present = false;
                            for (size_t t = static_cast<size_t>(first); t < out.index.size(); ++t)
                                present = present || out.index[t] == own.index[q];
                            if (!present) {
                                out.index.push_back(own.index[q]);
                                out.value.push_back(0.0);
                            }
                        }
                    }
        }
        out.count[c] = static_cast<int32_t>(out.index.size() - static_cast<size_t>(first));
    }
    return out;
}

// PhaseField::Write: int Nx, Ny, Nz, then per cell (x outer, z inner) size_t n and n x (size_t
// index, double value). Nx, Ny, Nz are the reference's Grid values (reduced axes 1).
inline bool write_phasefield(const std::string& path, const Dims& d, const Lists& lists)
{
    std::ofstream out(path, std::ios::out | std::ios::binary);
    if (!out) return false;
    const int n[3] = {d.nx, d.ny, d.nz};
    out.write(reinterpret_cast<const char*>(n), sizeof(n));
    size_t off = 0;
    for (size_t c = 0; c < lists.count.size(); ++c) {
        const size_t k = static_cast<size_t>(lists.count[c]);
        out.write(reinterpret_cast<const char*>(&k), sizeof(k));
        for (size_t q = 0; q < k; ++q, ++off) {
            const size_t idx = static_cast<size_t>(lists.index[off]);
            out.write(reinterpret_cast<const char*>(&idx), sizeof(idx));
            out.write(reinterpret_cast<const char*>(&lists.value[off]), sizeof(double));
        }
    }
    return static_cast<bool>(out);
}

// PhaseField::Read: the dimensions must match; returns false otherwise.
inline bool read_phasefield(const std::string& path, const Dims& d, Lists& lists)
{
    std::ifstream inp(path, std::ios::in | std::ios::binary);
    if (!inp) return false;
    int n[3] = {0, 0, 0};
    inp.read(reinterpret_cast<char*>(n), sizeof(n));
    if (n[0] != d.nx || n[1] != d.ny || n[2] != d.nz) return false;
    const long N = d.cells();
    lists.count.assign(N, 0);
    lists.index.clear();
    lists.value.clear();
    for (long c = 0; c < N; ++c) {
        size_t k = 0;
        inp.read(reinterpret_cast<char*>(&k), sizeof(k));
        lists.count[c] = static_cast<int32_t>(k);
        for (size_t q = 0; q < k; ++q) {
            size_t idx = 0;
            double v = 0.0;
            inp.read(reinterpret_cast<char*>(&idx), sizeof(idx));
            inp.read(reinterpret_cast<char*>(&v), sizeof(v));
            lists.index.push_back(static_cast<int32_t>(idx));
            lists.value.push_back(v);
        }
    }
    return static_cast<bool>(inp);
}

// One Grain record (include/GrainsProperties.h:51-78) with the fields this build keeps; the rest
// hold Grain::Clear's values (Mobile, zero kinematics and inertia).
struct GrainRecord {
    bool exist = true;
    bool mobile = true;
    int stage = 2;                 // GrainStages: 0 Seed, 1 Nucleus, 2 Stable
    size_t parent = 0;
    size_t phase = 0;
    size_t variant = 0;
    int state = 0;                 // AggregateStates: 0 Solid, 1 Liquid, 2 Gas
    double density = 0.0;
    double ref_volume = 1.0;
    double volume = 0.0;
    double max_volume = 0.0;
    double orientation[4] = {0.0, 0.0, 0.0, 0.0};  // quaternion as stored (s, x, y, z)
};

// GrainsProperties::Write: size_t n, then per grain the Grain::Write fields in order (bools as one
// byte, ints 4, size_t 8, doubles 8; the 3-vectors and the inertia matrix interleaved per row).
inline bool write_grains(const std::string& path, const std::vector<GrainRecord>& grains)
{
    std::ofstream out(path, std::ios::out | std::ios::binary);
    if (!out) return false;
    const size_t n = grains.size();
    out.write(reinterpret_cast<const char*>(&n), sizeof(n));
    const double zero = 0.0;
    for (const GrainRecord& g : grains) {
        out.write(reinterpret_cast<const char*>(&g.exist), sizeof(bool));
        out.write(reinterpret_cast<const char*>(&g.mobile), sizeof(bool));
        out.write(reinterpret_cast<const char*>(&g.stage), sizeof(int));
        out.write(reinterpret_cast<const char*>(&g.parent), sizeof(size_t));
        out.write(reinterpret_cast<const char*>(&g.phase), sizeof(size_t));
        out.write(reinterpret_cast<const char*>(&g.variant), sizeof(size_t));
        out.write(reinterpret_cast<const char*>(&g.state), sizeof(int));
        out.write(reinterpret_cast<const char*>(&g.density), sizeof(double));
        out.write(reinterpret_cast<const char*>(&g.ref_volume), sizeof(double));
        out.write(reinterpret_cast<const char*>(&g.volume), sizeof(double));
        out.write(reinterpret_cast<const char*>(&g.max_volume), sizeof(double));
        // Rcm, Vcm, Acm, aVel, aAcc, Force, Torque [i] and InertiaM(i, 0..2), i = 0..2
        for (int i = 0; i < 3 * (7 + 3); ++i) out.write(reinterpret_cast<const char*>(&zero), sizeof(double));
        out.write(reinterpret_cast<const char*>(g.orientation), 4 * sizeof(double));
    }
    return static_cast<bool>(out);
}

inline bool read_grains(const std::string& path, std::vector<GrainRecord>& grains)
{
    std::ifstream inp(path, std::ios::in | std::ios::binary);
    if (!inp) return false;
    size_t n = 0;
    inp.read(reinterpret_cast<char*>(&n), sizeof(n));
    grains.assign(n, GrainRecord{});
    double skip[3 * (7 + 3)];
    for (GrainRecord& g : grains) {
        inp.read(reinterpret_cast<char*>(&g.exist), sizeof(bool));
        inp.read(reinterpret_cast<char*>(&g.mobile), sizeof(bool));
        inp.read(reinterpret_cast<char*>(&g.stage), sizeof(int));
        inp.read(reinterpret_cast<char*>(&g.parent), sizeof(size_t));
        inp.read(reinterpret_cast<char*>(&g.phase), sizeof(size_t));
        inp.read(reinterpret_cast<char*>(&g.variant), sizeof(size_t));
        inp.read(reinterpret_cast<char*>(&g.state), sizeof(int));
        inp.read(reinterpret_cast<char*>(&g.density), sizeof(double));
        inp.read(reinterpret_cast<char*>(&g.ref_volume), sizeof(double));
        inp.read(reinterpret_cast<char*>(&g.volume), sizeof(double));
        inp.read(reinterpret_cast<char*>(&g.max_volume), sizeof(double));
        inp.read(reinterpret_cast<char*>(skip), sizeof(skip));
        inp.read(reinterpret_cast<char*>(g.orientation), 4 * sizeof(double));
    }
    return static_cast<bool>(inp);
}

// PhaseField::WriteVTK (single resolution, CurvatureOutput false): an ASCII structured grid with
// Interfaces, Flags, PhaseFields (majority index), PhaseFraction_n, Junctions (list length with the
// appended zero-valued entries), Variants, ParentGrain; `full` from full_lists. `fractions`
// ([phase * cells + cell], PhaseField::Fractions) when given, else summed here from `full`.
inline std::string vtk_text(const Dims& d, int nphases, const std::vector<GrainRecord>& grains, const Lists& full,
                            const std::vector<int8_t>& flag, int precision = 16,
                            const std::vector<double>* fractions = nullptr)
{
    const long N = d.cells();
    const std::vector<long> off = full.offsets();
    std::ostringstream b;
    b << "<?xml version= \"1.0\" encoding=\"UTF-8\" standalone=\"yes\"?>\n";
    b << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    b << "<StructuredGrid WholeExtent=\"" << 0 << " " << d.nx - 1 << " " << 0 << " " << d.ny - 1 << " " << 0 << " "
      << d.nz - 1 << "\">\n";
    b << "<PointData Scalars= \"ScalarData\">\n";
    auto header = [&](const std::string& name, const char* type) {
        b << "<DataArray type = \"" << type << "\" Name = \"" << name << "\" NumberOfComponents=\"" << 1
          << "\" format=\"" << "ascii" << "\">\n";
    };
    // points in VTK order: x fastest
    auto each = [&](auto fn) {
        for (int k = 0; k < d.nz; ++k)
            for (int j = 0; j < d.ny; ++j)
                for (int i = 0; i < d.nx; ++i) b << fn((static_cast<long>(i) * d.ny + j) * d.nz + k) << "\n";
        b << "</DataArray>\n";
    };
    auto as_double = [&]() { b << std::defaultfloat << std::setprecision(precision); };
    auto as_int = [&]() { b << std::fixed << std::setprecision(0); };
    // PhaseField::Interfaces (src/PhaseField.cpp:3765-3777)
    as_double();
    header("Interfaces", "Float64");
    each([&](long c) {
        double sum = 0.5;
        if (flag[c] == 2)
            for (long a = off[c]; a < off[c + 1]; ++a)
                for (long q = a + 1; q < off[c + 1]; ++q) sum -= full.value[a] * full.value[q];
        return 1.0 / (sum * 2.0);
    });
    as_int();
    header("Flags", "Int32");
    each([&](long c) { return static_cast<int>(flag[c]); });
    // NodePF::majority_index (include/Containers/NodePF.h:194-212)
    as_int();
    header("PhaseFields", "Int32");
    each([&](long c) {
        if (off[c + 1] - off[c] == 1) return static_cast<int>(full.index[off[c]]);
        double mv = 0.0;
        int mi = 0;
        for (long q = off[c]; q < off[c + 1]; ++q)
            if (full.value[q] > mv) { mv = full.value[q]; mi = full.index[q]; }
        return mi;
    });
    // PhaseField::CalculateFractions (src/PhaseField.cpp:411-423)
    for (int n = 0; n < nphases; ++n) {
        as_double();
        header("PhaseFraction_" + std::to_string(n), "Float64");
        each([&](long c) {
            if (fractions) return (*fractions)[static_cast<size_t>(n) * N + c];
            double f = 0.0;
            for (long q = off[c]; q < off[c + 1]; ++q)
                if (static_cast<int>(grains[full.index[q]].phase) == n) f += full.value[q];
            return f;
        });
    }
    as_int();
    header("Junctions", "UInt64");
    each([&](long c) { return static_cast<size_t>(off[c + 1] - off[c]); });
    // PhaseField::Variants / ParentGrain (src/PhaseField.cpp:3793-3839): of the largest entry
    for (int field = 0; field < 2; ++field) {
        as_double();
        header(field == 0 ? "Variants" : "ParentGrain", "Float64");
        each([&](long c) {
            double mv = 0.0;
            int r = 0;
            for (long q = off[c]; q < off[c + 1]; ++q)
                if (full.value[q] > mv) {
                    mv = full.value[q];
                    const GrainRecord& g = grains[full.index[q]];
                    r = static_cast<int>(field == 0 ? g.variant : g.parent);
                }
            return static_cast<double>(r);
        });
    }
    b << "</PointData>\n";
    b << "<Points>\n";
    b << "<DataArray type = \"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    b << std::fixed << std::setprecision(0);
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i)
                b << (i + 0) * 1.0 - 0.0 << " " << (j + 0) * 1.0 - 0.0 << " " << (k + 0) * 1.0 - 0.0 << "\n";
    b << "</DataArray>\n";
    b << "</Points>\n";
    b << "</StructuredGrid>\n";
    b << "</VTKFile>\n";
    return b.str();
}

inline bool write_vtk(const std::string& path, const Dims& d, int nphases, const std::vector<GrainRecord>& grains,
                      const Lists& full, const std::vector<int8_t>& flag, int precision = 16,
                      const std::vector<double>* fractions = nullptr)
{
    std::ofstream out(path);
    if (!out) return false;
    out << vtk_text(d, nphases, grains, full, flag, precision, fractions);
    return static_cast<bool>(out);
}

// FileInterface::MakeFileName (src/FileInterface.cpp:108-129): directory / (base + 8-digit step,
// when step >= 0, + extension)
inline std::string make_file_name(const std::string& dir, const std::string& base, int step, const std::string& ext)
{
    std::string name = base;
    if (step >= 0) {
        std::string num = std::to_string(step);
        if (num.size() < 8) num.insert(0, 8 - num.size(), '0');
        name += num;
    }
    name += ext;
    if (dir.empty()) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

}  // namespace io
}  // namespace openphase_amd
