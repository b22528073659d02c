// NmrRelaxationSolve.h -- drop-in replacement of the reference's host class
// (headers/NmrRelaxationSolve.h:32-130), same public surface:
//
//   NmrRelaxationSolve(int size, int rank, int* num);
//   int MPI_size, MPI_rank, PulseNum[3];
//   void init();       reads data/Nmr_start.dat, data/Nmr_data.dat,
//                      data/Nmr_grad.dat, data/Poro.dat, data/V[xyz].dat
//                      (NmrRelaxationSolve.cpp:20-33, 68-2534)
//   void solve();      the CPU-cadence loop of NmrRelaxationSolve.cpp:35-66
//                      (signal before each update, results_cpu.csv) with the
//                      upwind ExplicitSimple update (cpp:3148-3282)
//   void solve_gpu();  the GPU-cadence loop of NmrRelaxationSolveKernels.cu:1056-1319
//                      (update, pulses, signal; results_gpu.csv and
//                      results_NMR/results_grad_<g>_<g>_<g>.csv) with the
//                      MacCormack update (ExplicitMakUpgrade, cu:1473-1501)
//
// Both loops run the update on the MI355X through the C ABI of nmr_hip.h;
// there is no CPU solver behind them.  The whole domain lives on one GPU, so
// MPI_size must be 1.  Two-phase mixtures (NCOMP=2, data/n1_init.dat) and
// face-wise relaxivity files run with per-voxel coefficients and, for
// NCOMP=2, the 7-column two-phase output.  The random noise field
// (AMPL_NOISE with NDIRECT_FIELD 1-3, cpp:1973-1978) is drawn as the
// reference draws it (one rand() per padded cell from glibc's unseeded state)
// and runs as a per-voxel field.  Configurations the GPU path does not
// implement (NB flag combinations whose ghosts the reference never
// refreshes) are reported and end the program with exit(1), where the
// reference would run on with uninitialised data.  The NB flags select
// zero-derivative, periodic or zero-value ghosts in both loops (the
// reference's host loop ignores them).
#ifndef NMR_RELAXATION_SOLVE_H
#define NMR_RELAXATION_SOLVE_H

#include <cstdint>
#include <string>
#include <vector>

struct nmr_ctx;

class NmrRelaxationSolve {
public:
    NmrRelaxationSolve(int size, int rank, int* num)
    {
        MPI_size = size;
        MPI_rank = rank;
        PulseNum[0] = num[0];
        PulseNum[1] = num[1];
        PulseNum[2] = num[2];
    }
    ~NmrRelaxationSolve();
    int MPI_size, MPI_rank, PulseNum[3];
    void solve();
    void init();
    void solve_gpu();

    // Not in the reference: parse only the grid and Nmr_data.dat /
    // Nmr_grad.dat (no support checks, no field files), the parsed
    // configuration as JSON (for tests and tools), the directory the data/
    // files are read from, and the update scheme (-1: the reference's own,
    // MacCormack in solve_gpu() and the upwind ExplicitSimple in solve();
    // or NMR_SCHEME_UPWIND / NMR_SCHEME_MACCORMACK for both).
    void read_config();
    std::string config_json() const;
    void set_root(const std::string& dir) { root_ = dir; }
    void set_scheme(int scheme) { scheme_ = scheme; }

private:
    struct Pulse {
        int flag;
        double param, ax, ay, az;
    };

    // grid (GridSize gs of the reference)
    int ik = 0, jk = 0, kk = 0;
    // Nmr_data.dat keys
    int key = 0, iPrint = 0, jPrint = 0, kPrint = 0, ndPrint = 1, ndWrite = 0, ndVid = 0, ndAmpl = 0;
    int nPoro = 0, nComponents = 1, nFluid = 1, nSurfRelWater = 0, nSurfRelOil = 0;
    int nVectorMult = 0, nVelocity = 0, nDiffusion = 0, nDirectionField = 3, nSequence = 0;
    int nBoundary[6] = {1, 1, 1, 1, 1, 1};
    int bcMode[3] = {0, 0, 0}; // NMR_BC_* per axis from the NB flags
    int bufferSizeX = 0, bufferSizeY = 0, bufferSizeZ = 0;
    double dx = 0, dy = 0, dz = 0, bsx = 0, bsy = 0, bsz = 0, sequenceTime1 = 0, sequenceTime2 = 0;
    double constGradX = 0, constGradY = 0, constGradZ = 0, noiseAmpl = 0, velocityMultiplicator = 1;
    double dOil = 0, dWater = 0, r1Oil = 0, r1Water = 0, r2Oil = 0, r2Water = 0;
    double t1Oil = 0, t1Water = 0, t2Oil = 0, t2Water = 0, temperature = 0;
    double an1 = 0, an2 = 0, bn1 = 0, bn2 = 0, cfl = 0, cfl1 = 0, step0 = 0;
    // derived
    double gradientPulse = 0, multipleGradients[3] = {0, 0, 0}, direction[3] = {0, 0, 0};
    double endTime = 0, dTimeVel = 1, dTimeDiff = 1, dTimeRel = 1, dTimeNoise = 0, dTimeGrad = 0, dTime = 0;
    double equilibriumMagnetization = 0;
    int64_t nStep = 0;
    int flag1 = 0, flag2 = 0, flag3 = 0, flag4 = 0, flag5 = 0;
    // padded host arrays, index i + (ik+2)*(j + (jk+2)*k)
    std::vector<uint8_t> actCell;
    std::vector<double> vel[3], magn[3];
    std::vector<double> conc[2];   // concentration[0], [1] (cpp:930-1008)
    std::vector<double> surf[12];  // face relaxivities 1W xyz, 2W xyz, 1O xyz, 2O xyz (cpp:1010-1433)
    std::vector<double> dvox, r1vox, r2vox; // FillInDiffusinAndTime: D, 1/T1, 1/T2 per voxel
    bool perVoxel = false;         // mixture or face-wise relaxivity: per-voxel coefficients on the GPU
    bool initialised_ = false;
    int scheme_ = -1;
    std::string root_ = ".";
    nmr_ctx* ctx_ = nullptr;

    size_t index(int i, int j, int k) const { return (size_t)i + (size_t)(ik + 2) * ((size_t)j + (size_t)(jk + 2) * k); }
    std::string path(const char* rel) const { return root_ + "/" + rel; }
    void ReadGridSize();
    void ReadStart();
    void Validate() const;
    void ReadActiveCells();
    void ReadVelocity();
    void ReadConcentration();
    void ReadSurfRel();
    void Phase(size_t c, double* c1, double* c2) const;
    void Coefficients();
    void FillInMagnetization();
    void StepTime();
    void create_context(int scheme);
    std::vector<Pulse> SequenceGpu(double currentTime);
    std::vector<Pulse> SequenceCpu(double currentTime);
    void apply(const std::vector<Pulse>& pulses);
    void check(int rc, const char* what) const;
};

#endif
