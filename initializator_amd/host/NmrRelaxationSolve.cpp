// NmrRelaxationSolve.cpp -- host side of the drop-in class: the reference's
// input files, step-size rules, pulse sequences and output cadence, driving
// the gfx950 solver through the C ABI (include/nmr_hip.h).  Each function
// cites the reference code whose behaviour it keeps.
#include "NmrRelaxationSolve.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "nmr_hip.h"

namespace {

const double kG = 4.78941714e+7;  // cpp:807
const double kPi = 3.1415926535897932384626433832795028841971;
const double kSmall = 1.0e-50;     // cpp:809

double rmin(double a, double b) { return a <= b ? a : b; } // cpp:3289-3293

[[noreturn]] void unsupported(const std::string& what)
{
    std::cerr << "NmrRelaxationSolve: " << what << std::endl;
    std::exit(1);
}

// The C library's rand() as the reference calls it (glibc, never seeded:
// the state of srand(1)), restated so that the noise field does not depend
// on other rand() users in the process: glibc's TYPE_3 additive-feedback
// generator, r[0] = 1, r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31,
// r[i] = r[i-31] for i = 31..33, then r[i] = r[i-31] + r[i-3] mod 2^32, the
// first 310 of those discarded; each call returns r >> 1.
class GlibcRand {
public:
    GlibcRand()
    {
        int64_t w = 1;
        r_[0] = 1u;
        for (int i = 1; i < 31; i++) {
            w = (16807LL * w) % 2147483647LL;
            r_[i] = (uint32_t)w;
        }
        for (int i = 31; i < 34; i++) r_[i] = r_[i - 31];
        for (int k = 0; k < 310; k++) next();
    }
    int operator()() { return (int)(next() >> 1); }

private:
    uint32_t next()
    {
        const uint32_t v = r_[(i_ - 31) % 34] + r_[(i_ - 3) % 34];
        r_[i_ % 34] = v;
        i_ += 1;
        return v;
    }
    uint32_t r_[34];
    long i_ = 34;
};

} // namespace

NmrRelaxationSolve::~NmrRelaxationSolve()
{
    if (ctx_) nmr_hip_destroy(ctx_);
}

void NmrRelaxationSolve::check(int rc, const char* what) const
{
    // gpuAssert prints and exits (headers/NmrRelaxationSolve.h:14-27)
    if (rc != NMR_OK) {
        std::cout << "GPUassert:" << nmr_hip_strerror(rc) << " " << what << ": " << nmr_hip_last_error() << std::endl
                  << std::endl;
        std::exit(1);
    }
}

// cpp:68-93 (kk divided by the rank count: one rank here)
void NmrRelaxationSolve::ReadGridSize()
{
    std::ifstream f(path("data/Nmr_start.dat"));
    if (!f) unsupported("no file data/Nmr_start.dat");
    std::string rest;
    f >> ik;
    std::getline(f, rest);
    f >> jk;
    std::getline(f, rest);
    f >> kk;
    std::getline(f, rest);
    if (!f || ik < 1 || jk < 1 || kk < 1) unsupported("data/Nmr_start.dat: need three positive grid sizes");
    kk = kk / MPI_size;
}

// cpp:175-817: 61 "value<TAB>description" lines; every key is the first line
// whose description contains it, parsed with atoi/atof.
void NmrRelaxationSolve::ReadStart()
{
    std::ifstream f(path("data/Nmr_data.dat"));
    if (!f) unsupported("no file data/Nmr_data.dat");
    std::vector<std::pair<std::string, std::string>> keys;
    for (int i = 0; i < 61; ++i) {
        std::string a, b;
        std::getline(f, a, '\t');
        std::getline(f, b, '\n');
        keys.emplace_back(a, b);
    }
    auto find = [&](const char* k) -> const std::string* {
        for (const auto& it : keys)
            if (it.second.find(k) != std::string::npos) return &it.first;
        return nullptr;
    };
    auto gi = [&](const char* k, int& v) {
        if (const std::string* s = find(k)) v = std::atoi(s->c_str());
    };
    auto gd = [&](const char* k, double& v) {
        if (const std::string* s = find(k)) v = std::atof(s->c_str());
    };
    gi("KEY", key);
    gi("IPRINT", iPrint);
    gi("JPRINT", jPrint);
    gi("KPRINT", kPrint);
    gi("Printing on", ndPrint);
    gi("NDWRIT", ndWrite);
    gi("NDVID", ndVid);
    gi("NDAMPL", ndAmpl);
    gi("NPOR", nPoro);
    gi("NCOMP", nComponents);
    gi("NFLUID", nFluid);
    gi("nSurfRelWater", nSurfRelWater);
    gi("nSurfRelOil", nSurfRelOil);
    gi("NVECTOR", nVectorMult);
    gi("NVELOC", nVelocity);
    gi("NDIFFUS", nDiffusion);
    gi("NDIRECT_FIELD", nDirectionField);
    gi("NSEQUENCE", nSequence);
    gi("NB1", nBoundary[0]);
    gi("NB2", nBoundary[1]);
    gi("NB3", nBoundary[2]);
    gi("NB4", nBoundary[3]);
    gi("NB5", nBoundary[4]);
    gi("NB6", nBoundary[5]);
    gd("DX", dx);
    gd("DY", dy);
    gd("DZ", dz);
    gd("BX0", bsx);
    gd("BY0", bsy);
    gd("BZ0", bsz);
    gd("TIME1", sequenceTime1);
    gd("TIME2", sequenceTime2);
    gd("GRAD_X", constGradX);
    gd("GRAD_Y", constGradY);
    gd("GRAD_Z", constGradZ);
    gd("AMPL_NOISE", noiseAmpl);
    gd("VEL_MULT", velocityMultiplicator);
    gd("D_OIL", dOil);
    gd("D_WAT", dWater);
    gd("R1_OIL", r1Oil);
    gd("R1_WAT", r1Water);
    gd("R2_OIL", r2Oil);
    gd("R2_WAT", r2Water);
    gd("T1_OIL", t1Oil);
    gd("T1_WAT", t1Water);
    gd("T2_OIL", t2Oil);
    gd("T2_WAT", t2Water);
    gd("T0", temperature);
    gd("AN(1)", an1);
    gd("AN(2)", an2);
    gd("BN(1)", bn1);
    gd("BN(2)", bn2);
    gd("CFL", cfl);
    gd("CFL1", cfl1);
    gd("TIMEW", step0);
    gi("bufferX", bufferSizeX);
    gi("bufferY", bufferSizeY);
    gi("bufferZ", bufferSizeZ);

    // gradient table (cpp:727-778): line 1 is the count, entry n on line n+1
    std::ifstream g(path("data/Nmr_grad.dat"));
    if (!g) unsupported("no file data/Nmr_grad.dat");
    std::vector<double> table;
    int total = 0;
    g >> total;
    double v;
    while (g >> v) table.push_back(v);
    if (total < PulseNum[0]) unsupported("gradient index beyond data/Nmr_grad.dat");
    auto entry = [&](int n) -> double {
        if (n < 1 || n > (int)table.size()) return 0.0;
        return table[(size_t)n - 1];
    };
    gradientPulse = entry(PulseNum[0]);
    for (int d = 0; d < 3; d++) multipleGradients[d] = PulseNum[d] != 0 ? entry(PulseNum[d]) : 0.0;

    // cpp:781-810
    const double b_total = std::sqrt(bsx * bsx + bsy * bsy + bsz * bsz);
    if (nSequence == 0 || nSequence == 1) endTime = 2 * t2Water;
    else if (nSequence == 2) endTime = 4.0 * sequenceTime1 + sequenceTime2;
    else if (nSequence == 5 || nSequence == 9 || nSequence == 99 || nSequence == 98)
        endTime = sequenceTime2 + 2.0 * sequenceTime1;
    else endTime = 0.0; // the reference leaves it unset; Validate() refuses the run
    for (int i = 0; i < 3; i++) {
        direction[i] = 0.0;
        if (nDirectionField == i + 1) direction[i] = 1.0;
        if (nDirectionField == 0) direction[i] = (i == 0 ? bsx : i == 1 ? bsy : bsz) / b_total;
    }
    equilibriumMagnetization =
        1.0 / double((ik - 2 * bufferSizeX) * (jk - 2 * bufferSizeY) * (kk - 2 * bufferSizeZ));
    flag1 = flag2 = flag3 = flag4 = flag5 = 0;
}

// Ghost fill per axis from the NB flags, with the rules of
// BoundaryConditionsForVectorField_gpu (cu:1328-1396): x zero derivative when
// NB1 = NB3 = 1, periodic when both are 3, zero value when either is 2 (that
// rule runs last); y and z zero derivative or periodic when both flags agree.
// Other combinations leave the reference's ghosts stale: -1.
static void boundary_modes(const int nb[6], int out[3])
{
    out[0] = -1;
    if (nb[0] == 1 && nb[2] == 1) out[0] = NMR_BC_NEUMANN;
    if (nb[0] == 3 && nb[2] == 3) out[0] = NMR_BC_PERIODIC;
    if (nb[0] == 2 || nb[2] == 2) out[0] = NMR_BC_DIRICHLET;
    const int pairs[2][2] = {{1, 3}, {4, 5}};
    for (int a = 0; a < 2; a++) {
        const int lo = nb[pairs[a][0]], hi = nb[pairs[a][1]];
        out[a + 1] = (lo == 1 && hi == 1) ? NMR_BC_NEUMANN : ((lo == 3 && hi == 3) ? NMR_BC_PERIODIC : -1);
    }
}

// Inputs the GPU path implements (one or two fluids, any relaxivity files,
// the random noise field, boundary flags that define every ghost).
void NmrRelaxationSolve::Validate() const
{
    if (nComponents != 1 && nComponents != 2) unsupported("NCOMP must be 1 or 2");
    if (nComponents == 1 && nFluid != 1 && nFluid != 2) unsupported("NFLUID must be 1 or 2");
    int modes[3];
    boundary_modes(nBoundary, modes);
    for (int a = 0; a < 3; a++)
        if (modes[a] < 0)
            unsupported(std::string("NB flags leave the ") + "xyz"[a] +
                        " ghosts undefined (use 1 or 3 on both sides, or 2 on x)");
    if (ndPrint < 1) unsupported("'Printing on' must be >= 1");
    if (!(nSequence == 0 || nSequence == 1 || nSequence == 2 || nSequence == 5 || nSequence == 9 ||
          nSequence == 98 || nSequence == 99))
        unsupported("NSEQUENCE " + std::to_string(nSequence) + " is not a sequence of the reference");
}

void NmrRelaxationSolve::read_config()
{
    ReadGridSize();
    ReadStart();
}

// cpp:819-878: chars '0'/'1', k outer, j, i inner; all pores when NPOR=0.
// Ghosts copy the adjacent interior (cpp:880-928).
void NmrRelaxationSolve::ReadActiveCells()
{
    actCell.assign((size_t)(ik + 2) * (jk + 2) * (kk + 2), '0');
    if (nPoro == 0) {
        for (int k = 1; k <= kk; k++)
            for (int j = 1; j <= jk; j++)
                for (int i = 1; i <= ik; i++) actCell[index(i, j, k)] = '1';
    } else {
        std::ifstream f(path("data/Poro.dat"));
        if (!f) unsupported("no file data/Poro.dat");
        std::string header;
        std::getline(f, header);
        for (int k = 1; k <= kk; k++)
            for (int j = 1; j <= jk; j++)
                for (int i = 1; i <= ik; i++) {
                    char c = 0;
                    f >> c;
                    if (!f || (c != '0' && c != '1')) unsupported("data/Poro.dat: expected '0'/'1' cells");
                    actCell[index(i, j, k)] = (uint8_t)c;
                }
    }
    for (int k = 0; k < kk + 2; k++)
        for (int j = 0; j < jk + 2; j++) {
            actCell[index(0, j, k)] = actCell[index(1, j, k)];
            actCell[index(ik + 1, j, k)] = actCell[index(ik, j, k)];
        }
    for (int k = 0; k < kk + 2; k++)
        for (int i = 0; i < ik + 2; i++) {
            actCell[index(i, 0, k)] = actCell[index(i, 1, k)];
            actCell[index(i, jk + 1, k)] = actCell[index(i, jk, k)];
        }
    for (int j = 0; j < jk + 2; j++)
        for (int i = 0; i < ik + 2; i++) {
            actCell[index(i, j, 0)] = actCell[index(i, j, 1)];
            actCell[index(i, j, kk + 1)] = actCell[index(i, j, kk)];
        }
}

// cpp:1435-1591: staggered faces, i outer / k inner, times VEL_MULT;
// dTimeVel from the largest |raw value|.
void NmrRelaxationSolve::ReadVelocity()
{
    for (int d = 0; d < 3; d++) vel[d].assign(actCell.size(), 0.0);
    dTimeVel = 1.0;
    if (nVelocity != 1) return;
    double vmax = 0.0, vmin = 0.0;
    const char* names[3] = {"data/Vx.dat", "data/Vy.dat", "data/Vz.dat"};
    for (int d = 0; d < 3; d++) {
        std::ifstream f(path(names[d]));
        if (!f) unsupported(std::string("no file ") + names[d]);
        std::string header;
        std::getline(f, header);
        const int ni = ik + (d == 0), nj = jk + (d == 1), nk = kk + (d == 2);
        for (int i = 1; i <= ni; i++)
            for (int j = 1; j <= nj; j++)
                for (int k = 1; k <= nk; k++) {
                    double b = 0;
                    f >> b;
                    if (!f) unsupported(std::string(names[d]) + ": too few values");
                    vel[d][index(i, j, k)] = velocityMultiplicator * b;
                    if (vmax < b) vmax = b;
                    if (vmin > b) vmin = b;
                }
    }
    if (vmax > std::abs(vmin)) vmax = vmax * velocityMultiplicator;
    else vmax = std::abs(vmin) * velocityMultiplicator;
    if ((dx <= dy) && (dx <= dz)) dTimeVel = cfl * dx / (vmax + kSmall);
    if ((dy <= dy) && (dy <= dz)) dTimeVel = cfl * dy / (vmax + kSmall);
    if ((dz <= dy) && (dz <= dz)) dTimeVel = cfl * dz / (vmax + kSmall);
}

// Concentrations (cpp:930-1008): one fluid fills concentration[0] (water)
// or [1] (oil) with 1 on the interior; NCOMP=2 reads data/n1_init.dat
// (header, then i outer, j, k inner) into [0] and 1 - value into [1].
void NmrRelaxationSolve::ReadConcentration()
{
    for (int q = 0; q < 2; q++) conc[q].assign(actCell.size(), 0.0);
    if (nComponents == 1) {
        const int q = nFluid == 1 ? 0 : 1;
        for (int k = 1; k <= kk; k++)
            for (int j = 1; j <= jk; j++)
                for (int i = 1; i <= ik; i++) conc[q][index(i, j, k)] = 1.0;
        return;
    }
    std::ifstream f(path("data/n1_init.dat"));
    if (!f) unsupported("no file data/n1_init.dat (NCOMP=2)");
    std::string header;
    std::getline(f, header);
    for (int i = 1; i <= ik; i++)
        for (int j = 1; j <= jk; j++)
            for (int k = 1; k <= kk; k++) {
                float b = 0.f; // the reference reads into a float
                f >> b;
                if (!f) unsupported("data/n1_init.dat: too few values");
                conc[0][index(i, j, k)] = b;
                conc[1][index(i, j, k)] = 1.0 - b;
            }
}

// Face relaxivities (cpp:1010-1433): with nSurfRelWater / nSurfRelOil = 1
// the files data/SurfRel{1,2}{Water,Oil}{X,Y,Z}.dat (header, i outer, k
// inner, staggered like the velocity files); otherwise the uniform
// R1/R2 of that fluid on every face.
void NmrRelaxationSolve::ReadSurfRel()
{
    const char* fluid[2] = {"Water", "Oil"};
    const int on[2] = {nSurfRelWater, nSurfRelOil};
    const double uni[2][2] = {{r1Water, r2Water}, {r1Oil, r2Oil}};
    for (int fl = 0; fl < 2; fl++)
        for (int rel = 0; rel < 2; rel++)
            for (int ax = 0; ax < 3; ax++) {
                std::vector<double>& a = surf[fl * 6 + rel * 3 + ax];
                if (on[fl] != 1) {
                    a.assign(actCell.size(), uni[fl][rel]);
                    continue;
                }
                a.assign(actCell.size(), 0.0);
                const std::string name = std::string("data/SurfRel") + (rel == 0 ? "1" : "2") + fluid[fl] + "XYZ"[ax] +
                                         ".dat";
                std::ifstream f(path(name.c_str()));
                if (!f) unsupported("no file " + name);
                std::string header;
                std::getline(f, header);
                const int ni = ik + (ax == 0), nj = jk + (ax == 1), nk = kk + (ax == 2);
                for (int i = 1; i <= ni; i++)
                    for (int j = 1; j <= nj; j++)
                        for (int k = 1; k <= nk; k++) {
                            double b = 0;
                            f >> b;
                            if (!f) unsupported(name + ": too few values");
                            a[index(i, j, k)] = b;
                        }
            }
}

// Phase fractions of cell c (cpp:1780-1791).
void NmrRelaxationSolve::Phase(size_t c, double* c1, double* c2) const
{
    if (nComponents == 2) {
        const double x0 = conc[0][c], x1 = conc[1][c];
        const double z1 = std::sqrt((x0 - an1) * (x0 - an1) + (x1 - an2) * (x1 - an2));
        const double z2 = std::sqrt((x0 - bn1) * (x0 - bn1) + (x1 - bn2) * (x1 - bn2));
        *c1 = z2 / (z1 + z2);
        *c2 = z1 / (z1 + z2);
    } else {
        *c1 = conc[0][c];
        *c2 = conc[1][c];
    }
}

// FillInDiffusinAndTime (cpp:1756-1864) and the step limits of
// FillInMainField (cpp:1980-1988): per-voxel D, 1/T1, 1/T2 (kept for the GPU
// when the coefficients vary), their extremes for StepTime.
void NmrRelaxationSolve::Coefficients()
{
    double dMax = 0.0, tMin = 100.0;
    dvox.assign(actCell.size(), 0.0);
    r1vox.assign(actCell.size(), 0.0);
    r2vox.assign(actCell.size(), 0.0);
    const double h3[3] = {dx, dy, dz};
    for (int k = 1; k <= kk; k++)
        for (int j = 1; j <= jk; j++)
            for (int i = 1; i <= ik; i++) {
                const size_t c = index(i, j, k);
                if (actCell[c] != '1') continue;
                double c1, c2;
                Phase(c, &c1, &c2);
                double D = (c1 * dWater + c2 * dOil) * std::pow(1.0 - 2.0 * c1 * c2 / (c1 * c1 + c2 * c2), 10.0);
                if (nDiffusion == 0) D = kSmall;
                double ct1 = c1 / t1Water + c2 / t1Oil, ct2 = c1 / t2Water + c2 / t2Oil;
                const size_t nb[6] = {index(i - 1, j, k), index(i + 1, j, k), index(i, j - 1, k),
                                      index(i, j + 1, k), index(i, j, k - 1), index(i, j, k + 1)};
                const size_t fc[6] = {c, nb[1], c, nb[3], c, nb[5]};
                for (int s = 0; s < 6; s++)
                    if (actCell[nb[s]] == '0') {
                        const int ax = s / 2;
                        ct1 += (c1 * surf[ax][fc[s]] + c2 * surf[6 + ax][fc[s]]) / h3[ax];
                        ct2 += (c1 * surf[3 + ax][fc[s]] + c2 * surf[9 + ax][fc[s]]) / h3[ax];
                    }
                dvox[c] = nDiffusion == 1 ? D : 0.0;
                r1vox[c] = ct1;
                r2vox[c] = ct2;
                const double t1 = 1.0 / ct1, t2 = 1.0 / ct2;
                if (dMax <= D) dMax = D;
                if (tMin >= t1) tMin = t1;
                if (tMin >= t2) tMin = t2;
            }
    if (nDiffusion == 1) {
        if ((dx <= dy) && (dx <= dz)) dTimeDiff = 2.0 * cfl * dx * dx / dMax;
        if ((dy <= dy) && (dy <= dz)) dTimeDiff = 2.0 * cfl * dy * dy / dMax;
        if ((dz <= dy) && (dz <= dz)) dTimeDiff = 2.0 * cfl * dz * dz / dMax;
    } else {
        dTimeDiff = 1.0;
    }
    dTimeRel = cfl1 * tMin;
    double gradAmpl = kSmall;
    if (constGradX != 0) gradAmpl = std::abs(constGradX * dx * ik);
    if (constGradY != 0) gradAmpl = std::abs(constGradY * dy * jk);
    if (constGradZ != 0) gradAmpl = std::abs(constGradZ * dz * kk);
    dTimeNoise = cfl1 / (kG * noiseAmpl);
    dTimeGrad = cfl1 / (kG * gradAmpl);
    perVoxel = nComponents == 2 || nSurfRelWater == 1 || nSurfRelOil == 1;
}

// cpp:2472-2517: pEq along the field direction in pores (ghosts included).
void NmrRelaxationSolve::FillInMagnetization()
{
    for (int d = 0; d < 3; d++) magn[d].assign(actCell.size(), 0.0);
    const double pEq = equilibriumMagnetization;
    const double bt = std::sqrt(bsx * bsx + bsy * bsy + bsz * bsz);
    for (size_t c = 0; c < actCell.size(); c++) {
        if (actCell[c] != '1') continue;
        if (nDirectionField == 1) magn[0][c] = pEq;
        if (nDirectionField == 2) magn[1][c] = pEq;
        if (nDirectionField == 3) magn[2][c] = pEq;
        if (nDirectionField == 0) {
            magn[0][c] = pEq * bsx / bt;
            magn[1][c] = pEq * bsy / bt;
            magn[2][c] = pEq * bsz / bt;
        }
    }
}

// cpp:2519-2534 (the step count is kept in 64 bits: the reference's int
// overflows for FID/CPMG end times of 2*T2 = 5e10 s and then runs no step).
void NmrRelaxationSolve::StepTime()
{
    if (sequenceTime1 != 0)
        dTime = rmin(rmin(rmin(dTimeNoise, dTimeGrad), rmin(dTimeVel, dTimeDiff)), rmin(dTimeRel, sequenceTime1 / 100.0));
    else
        dTime = rmin(rmin(rmin(dTimeNoise, dTimeGrad), rmin(dTimeVel, dTimeDiff)), rmin(dTimeRel, sequenceTime2 / 100.0));
    const double n = endTime / dTime;
    if (!(n >= 0.0) || n > 4.0e18) unsupported("step count endTime/dTime out of range");
    nStep = (int64_t)n + 1;
}

void NmrRelaxationSolve::init()
{
    if (MPI_size != 1 || MPI_rank != 0)
        unsupported("the whole domain is solved on one GPU: run one process (MPI_size 1)");
    ReadGridSize();
    ReadStart();
    Validate();
    boundary_modes(nBoundary, bcMode);
    ReadActiveCells();
    ReadVelocity();
    ReadConcentration();
    ReadSurfRel();
    Coefficients();
    FillInMagnetization();
    StepTime();
    initialised_ = true;
    std::cout << "dTime=" << std::scientific << dTime << " nStep=" << nStep << std::endl;
}

std::string NmrRelaxationSolve::config_json() const
{
    std::ostringstream o;
    o << std::setprecision(17);
    o << "{\"ik\": " << ik << ", \"jk\": " << jk << ", \"kk\": " << kk << ", \"ndPrint\": " << ndPrint
      << ", \"nPoro\": " << nPoro << ", \"nComponents\": " << nComponents << ", \"nFluid\": " << nFluid
      << ", \"nVectorMult\": " << nVectorMult << ", \"nVelocity\": " << nVelocity << ", \"nDiffusion\": " << nDiffusion
      << ", \"nDirectionField\": " << nDirectionField << ", \"nSequence\": " << nSequence << ", \"dx\": " << dx
      << ", \"dy\": " << dy << ", \"dz\": " << dz << ", \"bsx\": " << bsx << ", \"bsy\": " << bsy << ", \"bsz\": " << bsz
      << ", \"time1\": " << sequenceTime1 << ", \"time2\": " << sequenceTime2 << ", \"gradX\": " << constGradX
      << ", \"gradY\": " << constGradY << ", \"gradZ\": " << constGradZ << ", \"velMult\": " << velocityMultiplicator
      << ", \"dOil\": " << dOil << ", \"dWater\": " << dWater << ", \"r1Water\": " << r1Water << ", \"r2Water\": " << r2Water
      << ", \"t1Water\": " << t1Water << ", \"t2Water\": " << t2Water << ", \"cfl\": " << cfl << ", \"cfl1\": " << cfl1
      << ", \"bufferX\": " << bufferSizeX << ", \"bufferY\": " << bufferSizeY << ", \"bufferZ\": " << bufferSizeZ
      << ", \"gradientPulse\": " << gradientPulse << ", \"multipleGradients\": [" << multipleGradients[0] << ", "
      << multipleGradients[1] << ", " << multipleGradients[2] << "], \"endTime\": " << endTime
      << ", \"pEq\": " << equilibriumMagnetization << ", \"dTime\": " << dTime << ", \"nStep\": " << nStep
      << ", \"dTimeVel\": " << dTimeVel << ", \"dTimeDiff\": " << dTimeDiff << ", \"dTimeRel\": " << dTimeRel
      << ", \"nBoundary\": [" << nBoundary[0] << ", " << nBoundary[1] << ", " << nBoundary[2] << ", " << nBoundary[3]
      << ", " << nBoundary[4] << ", " << nBoundary[5] << "], \"bc\": [" << bcMode[0] << ", " << bcMode[1] << ", "
      << bcMode[2] << "]}";
    return o.str();
}

void NmrRelaxationSolve::create_context(int scheme)
{
    if (!initialised_) unsupported("init() must run before solve");
    if (ctx_) {
        nmr_hip_destroy(ctx_);
        ctx_ = nullptr;
    }
    const double c1 = nFluid == 1 ? 1.0 : 0.0, c2 = nFluid == 2 ? 1.0 : 0.0;
    nmr_problem p{};
    p.nx = ik;
    p.ny = jk;
    p.nz = kk;
    p.ncomp = 3;
    p.flags = (nDiffusion == 1 ? NMR_F_DIFFUSION : 0) | (nVelocity == 1 ? NMR_F_VELOCITY : 0) |
              (nVectorMult == 1 ? NMR_F_PRECESSION : 0);
    p.scheme = scheme_ >= 0 ? scheme_ : scheme;
    p.phase_mode = NMR_PHASE_LINEAR;
    p.bc[0] = bcMode[0];
    p.bc[1] = bcMode[1];
    p.bc[2] = bcMode[2];
    p.buffer[0] = bufferSizeX;
    p.buffer[1] = bufferSizeY;
    p.buffer[2] = bufferSizeZ;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dTime;
    p.diffusivity = c1 * dWater + c2 * dOil;
    p.r1_bulk = c1 / t1Water + c2 / t1Oil;
    p.r2_bulk = c1 / t2Water + c2 / t2Oil;
    p.rho1 = c1 * r1Water + c2 * r1Oil;
    p.rho2 = c1 * r2Water + c2 * r2Oil;
    p.gamma = kG;
    p.p_eq = equilibriumMagnetization;
    // mainField (cpp:1944-1964)
    p.b0_dir[0] = (nDirectionField == 1 || nDirectionField == 0) ? bsx : 0.0;
    p.b0_dir[1] = (nDirectionField == 2 || nDirectionField == 0) ? bsy : 0.0;
    p.b0_dir[2] = (nDirectionField == 3 || nDirectionField == 0) ? bsz : 0.0;
    p.grad[0] = constGradX;
    p.grad[1] = constGradY;
    p.grad[2] = constGradZ;
    check(nmr_hip_create(&p, &ctx_), "nmr_hip_create");
    check(nmr_hip_upload_mask(ctx_, actCell.data(), NMR_LAYOUT_PADDED), "nmr_hip_upload_mask");
    if (nVelocity == 1)
        check(nmr_hip_upload_velocity(ctx_, vel[0].data(), vel[1].data(), vel[2].data(), NMR_LAYOUT_PADDED),
              "nmr_hip_upload_velocity");
    check(nmr_hip_upload_magnetization(ctx_, magn[0].data(), magn[1].data(), magn[2].data(), NMR_LAYOUT_PADDED),
          "nmr_hip_upload_magnetization");
    if (perVoxel)
        check(nmr_hip_upload_coefficients(ctx_, dvox.data(), r1vox.data(), r2vox.data(), NMR_LAYOUT_PADDED),
              "nmr_hip_upload_coefficients");
    if (noiseAmpl != 0.0 && nDirectionField != 0) {
        // FillInMainField's noise (cpp:1973-1978): one rand() per padded cell,
        // k outermost, on the main-field axis, in place of its gradient field
        GlibcRand rnd;
        std::vector<double> nf(actCell.size());
        for (int k = 0; k < kk + 2; k++)
            for (int j = 0; j < jk + 2; j++)
                for (int i = 0; i < ik + 2; i++) nf[index(i, j, k)] = noiseAmpl * (rnd() % 2000 + 1 - 1000) / 1000;
        check(nmr_hip_upload_field(ctx_, nf.data(), NMR_LAYOUT_PADDED), "nmr_hip_upload_field");
    }
    if (nComponents == 2) { // two-phase outputs: weights c1, c2 on cells with c1 or c2 above 0.8
        std::vector<double> w1(actCell.size(), 0.0), w2(actCell.size(), 0.0);
        for (size_t c = 0; c < actCell.size(); c++) {
            if (actCell[c] != '1') continue;
            double c1, c2;
            Phase(c, &c1, &c2);
            if (c1 > 0.8 || c2 > 0.8) {
                w1[c] = c1;
                w2[c] = c2;
            }
        }
        check(nmr_hip_upload_phase_weights(ctx_, w1.data(), w2.data(), NMR_LAYOUT_PADDED),
              "nmr_hip_upload_phase_weights");
    }
}

void NmrRelaxationSolve::apply(const std::vector<Pulse>& pulses)
{
    for (const Pulse& q : pulses) check(nmr_hip_pulse(ctx_, q.flag, q.param, q.ax, q.ay, q.az), "nmr_hip_pulse");
}

// Pulse events of Sequence_gpu (cpp:2681-3146) at one time.
std::vector<NmrRelaxationSolve::Pulse> NmrRelaxationSolve::SequenceGpu(double t)
{
    std::vector<Pulse> out;
    const double* d = direction;
    auto win = [&](double t0) { return (t < t0 + dTime) && (t >= t0); };
    const double T1 = sequenceTime1, T2 = sequenceTime2;
    if (nSequence == 0) {
        if (win(0.0)) out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]});
    } else if (nSequence == 1) {
        if (win(0.0)) out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]});
        for (int i = 0; i < 20; i++)
            if (win((2 * i + 1) * T1)) out.push_back({2, kPi, d[2], d[0], d[1]});
    } else if (nSequence == 2) {
        if (win(0.0) && flag1 == 0) { out.push_back({2, 0.5 * kPi, -d[2], d[0], d[1]}); flag1 = 1; }
        if (win(T1) && flag2 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag2 = 1; }
        if (win(2.0 * T1) && flag3 == 0) { out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]}); flag3 = 1; }
        if (win(2.0 * T1 + T2) && flag4 == 0) { out.push_back({2, 0.5 * kPi, -d[2], d[0], d[1]}); flag4 = 1; }
        if (win(3.0 * T1 + T2) && flag5 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag5 = 1; }
        if (win(3.5 * T1 + T2) && (!flag1 || !flag2 || !flag3 || !flag4 || !flag5)) std::cout << "Sequence fail" << std::endl;
    } else if (nSequence == 9) {
        if (win(0.0) && flag1 == 0) { out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]}); flag1 = 1; }
        if (win(T1) && flag2 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag2 = 1; }
        if (win(T1 + T2 / 2) && flag3 == 0) { out.push_back({2, kPi, d[2], d[0], d[1]}); flag3 = 1; }
        if (win(T2 + T1) && flag4 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag4 = 1; }
    } else if (nSequence == 5) {
        if (win(0.0)) out.push_back({2, kPi, d[2], d[0], d[1]});
    } else if (nSequence == 99) {
        const double* m = multipleGradients;
        if (win(0.0) && flag1 == 0) flag1 = 1;
        if (win(T1) && flag2 == 0) { out.push_back({3, 1, m[0], m[1], m[2]}); flag2 = 1; }
        if (win(T1 + T2 / 2) && flag3 == 0) flag3 = 1;
        if (win(T2 + T1) && flag4 == 0) { out.push_back({3, 1, -m[0], -m[1], -m[2]}); flag4 = 1; }
    } else if (nSequence == 98) {
        const double* m = multipleGradients;
        if (win(0.0) && flag1 == 0) { out.push_back({2, 0.5 * kPi, 0, -1, 0}); flag1 = 1; }
        if (win(T1) && flag2 == 0) { out.push_back({3, 1, m[0], m[1], m[2]}); flag2 = 1; }
        if (win(T1 + T2 / 2) && flag3 == 0) { out.push_back({2, kPi, 0, 0, 1}); flag3 = 1; }
        if (win(T2 + T1) && flag4 == 0) { out.push_back({3, 1, m[0], m[1], m[2]}); flag4 = 1; }
    }
    return out;
}

// Pulse events of the CPU Sequence (cpp:2610-2679): FID, CPMG (100 echoes), PFG.
std::vector<NmrRelaxationSolve::Pulse> NmrRelaxationSolve::SequenceCpu(double t)
{
    std::vector<Pulse> out;
    const double* d = direction;
    auto win = [&](double t0) { return (t < t0 + dTime) && (t >= t0); };
    const double T1 = sequenceTime1, T2 = sequenceTime2;
    if (nSequence == 0) {
        if (win(0.0)) out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]});
    } else if (nSequence == 1) {
        if (win(0.0)) out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]});
        for (int i = 0; i < 100; i++)
            if (win((2 * i + 1) * T1)) out.push_back({2, kPi, d[2], d[0], d[1]});
    } else if (nSequence == 2) {
        if (win(0.0) && flag1 == 0) { out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]}); flag1 = 1; }
        if (win(T1) && flag2 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag2 = 1; }
        if (win(2.0 * T1) && flag3 == 0) { out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]}); flag3 = 1; }
        if (win(2.0 * T1 + T2) && flag4 == 0) { out.push_back({2, 0.5 * kPi, d[2], d[0], d[1]}); flag4 = 1; }
        if (win(3.0 * T1 + T2) && flag5 == 0) { out.push_back({1, gradientPulse, d[0], d[1], d[2]}); flag5 = 1; }
        if (win(3.5 * T1 + T2) && (!flag1 || !flag2 || !flag3 || !flag4 || !flag5)) std::cout << "Sequence fail" << std::endl;
    }
    return out;
}

namespace {

// One CSV row: Time; SumX; SumY; SumZ, or with two phases
// Time; SumX1; SumY1; SumZ1; SumX2; SumY2; SumZ2 (cu:1212-1218).
void csv_row(std::ofstream& f, double t, const double* s, int n)
{
    f << std::scientific << t;
    for (int q = 0; q < n; q++) f << "; " << std::scientific << s[q];
    f << std::endl;
}

void print_row(double t, const double* s, int n)
{
    if (n == 3) {
        std::cout << " Time = " << t << " SumX=" << s[0] << " SumY=" << s[1] << " SumZ=" << s[2] << std::endl;
    } else {
        std::cout << " Time = " << t << " SumX1=" << s[0] << " SumY1=" << s[1] << " SumZ1=" << s[2] << "\t"
                  << " SumX2=" << s[3] << " SumY2=" << s[4] << " SumZ2=" << s[5] << std::endl;
    }
}

} // namespace

// GPU cadence (NmrRelaxationSolveKernels.cu:1187-1227): for i = 0..nStep:
// update, pulses of Sequence_gpu at t = i*dt, then the signal row when
// i % ndPrint == 0 (the row stamped t = i*dt holds i+1 updates).
void NmrRelaxationSolve::solve_gpu()
{
    create_context(NMR_SCHEME_MACCORMACK);
    flag1 = flag2 = flag3 = flag4 = flag5 = 0;
    std::ofstream res(path("results_gpu.csv"));
    res << "Time; SumX; SumY; SumZ" << std::endl;
    std::vector<double> sig;
    const int ns = nComponents == 2 ? 6 : 3; // values per signal row
    double s[6] = {0, 0, 0, 0, 0, 0};
    double currentTime = 0.0;
    int64_t pending = 0; // updates not yet run
    for (int64_t i = 0; i < nStep + 1; i++) {
        currentTime = (double)i * dTime;
        pending += 1;
        std::vector<Pulse> pulses = SequenceGpu(currentTime);
        const bool out = (i % ndPrint) == 0;
        if (pulses.empty() && !out && i < nStep) continue;
        sig.resize((size_t)(pending + 1) * ns);
        check(nmr_hip_run(ctx_, pending, nullptr, sig.data()), "nmr_hip_run");
        pending = 0;
        if (!pulses.empty()) apply(pulses);
        if (out) {
            if (pulses.empty()) {
                for (int q = 0; q < ns; q++) s[q] = sig[sig.size() - ns + q];
            } else {
                check(nmr_hip_signal(ctx_, s), "nmr_hip_signal");
            }
            print_row(currentTime, s, ns);
            csv_row(res, currentTime, s, ns);
        }
    }
    res.close();
    // final point (cu:1290-1315)
    char name[256];
    std::snprintf(name, sizeof(name), "results_NMR/results_grad_%d_%d_%d.csv", PulseNum[0], PulseNum[1], PulseNum[2]);
    std::ofstream grad(path(name));
    check(nmr_hip_signal(ctx_, s), "nmr_hip_signal");
    print_row(currentTime, s, ns);
    if (grad) csv_row(grad, currentTime, s, ns);
    nmr_hip_destroy(ctx_);
    ctx_ = nullptr;
}

// CPU cadence (cpp:35-66): for i = 0..nStep: signal row when i % ndPrint
// == 0 (before this step's update), update, pulses of Sequence at t = i*dt.
void NmrRelaxationSolve::solve()
{
    create_context(NMR_SCHEME_UPWIND);
    flag1 = flag2 = flag3 = flag4 = flag5 = 0;
    std::ofstream res(path("results_cpu.csv"));
    res << "Time; SumX; SumY; SumZ" << std::endl;
    const int ns = nComponents == 2 ? 6 : 3;
    double s[6];
    int64_t i = 0;
    while (i < nStep + 1) {
        const double currentTime = (double)i * dTime;
        if ((i % ndPrint) == 0) {
            check(nmr_hip_signal(ctx_, s), "nmr_hip_signal");
            print_row(currentTime, s, ns);
            csv_row(res, currentTime, s, ns);
        }
        // updates i .. j-1 in one call: stop after a step that is followed by
        // a pulse, or before a step that starts with an output row
        int64_t j = i;
        std::vector<Pulse> pulses;
        for (;;) {
            pulses = SequenceCpu((double)j * dTime);
            j += 1;
            if (!pulses.empty() || j >= nStep + 1 || (j % ndPrint) == 0) break;
        }
        check(nmr_hip_run(ctx_, j - i, nullptr, nullptr), "nmr_hip_run");
        if (!pulses.empty()) apply(pulses);
        i = j;
    }
    res.close();
    nmr_hip_destroy(ctx_);
    ctx_ = nullptr;
}
