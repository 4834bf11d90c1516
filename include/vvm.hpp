// vvm.hpp -- C++ state and config API of the 2D vector vorticity model,
// with the timestep running on an MI355X through the C ABI of vvm_dev.h.
//
// The class and member names follow the reference's src/Declare.hpp so that
// code written against `Config_VVM`, `vvm` and `vvm::Iteration::TimeMarching`
// compiles against this header unchanged for the hot path:
//   - Config_VVM keeps the reference constructor (src/Declare.hpp:29-38);
//   - vvm keeps the host arrays `[i][k]` (k fastest) with the same names,
//     the contiguous `<name>cont` blocks (src/Declare.hpp:400-407) and the
//     AB2 `***` histories [i][k][2] (src/Declare.hpp:417-427);
//   - the static operator classes (src/Declare.hpp:707-941) keep their
//     names; each runs on the device and leaves the host arrays updated.
//
// Ownership is the reference's: vvm owns every host array.  The device
// handle (`dev`) belongs to vvm and is created by
// PoissonSolver::InitPoissonMatrix (or lazily by the first operator);
// device memory is released with the model.
//
// Errors: a device error prints the C layer's message and calls exit(1),
// as the reference's GPU back-end does (src/Declare.hpp:767-772).  A solve
// that misses its stop rule prints a warning and the step continues
// (src/PoissonSolver.cpp:463-467).
#pragma once

#include <functional>
#include <string>

#include "vvm_dev.h"

class Config_VVM {
   public:
    Config_VVM(double dt, double dx, double dz, int XRANGE, int ZRANGE, double TIMEEND, int TIMEROUTPUTSIZE,
               std::string outputpath, int OUTPUTSTEP, double Kx, double Kz, double TIMETS, double tolerance,
               double GRAVITY, double Cp, double Cv, double Rd, double Lv, double P0, double PSURF,
               double addforcingtime, int CASE, double mositure_nudge_time, int year, int month, int day, double hour,
               double minute, double second, double lon, double lat);
    ~Config_VVM() {}

    double dt, dx, dz;
    int XRANGE, ZRANGE;  // domain size in m, including one ghost cell each side
    double TIMEEND;
    int TIMEROUTPUTSIZE;
    std::string outputpath;
    int OUTPUTSTEP;
    double Kx, Kz, TIMETS, tolerance;
    double GRAVITY, Cp, Cv, Rd, Lv, P0, PSURF;
    double addforcingtime;
    int CASE;
    double mositure_nudge_time;
    int year, month, day;
    double hour, minute, second, lon, lat;

    // Device back-end options (no counterpart in the reference, which picks
    // its back-end at compile time).  Defaults reproduce the reference path:
    // FP64 solve from a zero guess, iterated to round-off.
    int gpu_id = 0;
    int solver_mode = VVM_SOLVER_FP64;
    int guess_mode = VVM_GUESS_ZERO;
    int stop_mode = VVM_STOP_PARITY;
    int max_iter = 2000;
    double tol_rel = 1e-10;
    int tropical_forcing = 0;
};

class vvm {
   public:
    explicit vvm(const Config_VVM &config);
    ~vvm();
    vvm(const vvm &) = delete;
    vvm &operator=(const vvm &) = delete;

    static double **allocate2DContinuousArray(int rows, int cols, double *&contMemory);
    static void deallocate2DContinuousArray(double **array, double *contMemory);
    double ***allocate3DContinuousArray(int dim1, int dim2, int dim3, double *&contMemory);
    void deallocate3DContinuousArray(double ***array, double *contMemory);

    // grid and constants (src/Declare.hpp:431-480)
    double rdx = 0, r2dx = 0, rdz = 0, r2dz = 0, rdx2 = 0, rdz2 = 0;
    int nx = 0, nz = 0;
    double dt = 0, d2t = 0, dx = 0, dz = 0;
    int XRANGE = 0, ZRANGE = 0;
    double TIMEEND = 0;
    int TIMEROUTPUTSIZE = 0;
    std::string outputpath;
    int OUTPUTSTEP = 0;
    double Kx = 0, Kz = 0, TIMETS = 0, tolerance = 0;
    double GRAVITY = 0, Cp = 0, Cv = 0, Rd = 0, Lv = 0, P0 = 0, PSURF = 0;
    double addforcingtime = 0;
    int CASE = 0;
    int year = 2025, month = 3, day = 20;
    double hour = 12., minute = 0., second = 0., lon = 0., lat = 0.;

    // 0-D state (src/Declare.hpp:482-487)
    int step = 0;
    double ubarTopp = 0, ubarTop = 0, ubarTopm = 0;
    double moisture_nudge_time = 0.;
    double dubarTop_advect[2] = {0, 0};

    // 1-D profiles, nz long (src/Declare.hpp:490-509)
    double *thb = nullptr, *thb_init = nullptr, *thbm = nullptr, *thb_zeta = nullptr;
    double *rhou = nullptr, *rhow = nullptr, *pib = nullptr, *qvb = nullptr, *qvb0 = nullptr, *qvsb = nullptr;
    double *pb = nullptr, *thvb = nullptr, *thvbm = nullptr, *z = nullptr, *z_zeta = nullptr, *lambda2 = nullptr;
    double *Q1LS = nullptr, *Q2LS = nullptr;
    // nx-long rows
    double *xi = nullptr, *uxi = nullptr, *precip = nullptr;

    // 2-D fields [i][k] (src/Declare.hpp:546-648)
    double **zetap = nullptr, **zeta = nullptr, **zetam = nullptr;
    double **thp = nullptr, **th = nullptr, **thm = nullptr;
    double **qvp = nullptr, **qv = nullptr, **qvm = nullptr;
    double **qcp = nullptr, **qc = nullptr, **qcm = nullptr;
    double **qrp = nullptr, **qr = nullptr, **qrm = nullptr;
    double **u = nullptr, **w = nullptr, **RKM = nullptr, **RKH = nullptr;
    double **evaporation = nullptr, **accretion = nullptr, **autoconversion = nullptr, **condensation = nullptr;
    double **init_th_forcing = nullptr;
    // radiation heating rate [K/day] that AddForcing adds to thp
    // (src/Declare.hpp:561, AddForcing.cpp:20); zero unless a radiation
    // scheme fills it (solve_radiation below)
    double **radiation_heating_rate = nullptr;
    double *radiation_heating_ratecont = nullptr;
    double *zetapcont = nullptr, *zetacont = nullptr, *zetamcont = nullptr;
    double *thpcont = nullptr, *thcont = nullptr, *thmcont = nullptr;
    double *qvpcont = nullptr, *qvcont = nullptr, *qvmcont = nullptr;
    double *qcpcont = nullptr, *qccont = nullptr, *qcmcont = nullptr;
    double *qrpcont = nullptr, *qrcont = nullptr, *qrmcont = nullptr;
    double *ucont = nullptr, *wcont = nullptr, *RKMcont = nullptr, *RKHcont = nullptr;
    double *evaporationcont = nullptr, *accretioncont = nullptr, *autoconversioncont = nullptr,
           *condensationcont = nullptr;
    double *init_th_forcingcont = nullptr;
    bool status_for_adding_forcing = true;

    // AB2 histories [i][k][2] (src/Declare.hpp:652-692)
    double ***dth_advect = nullptr, ***dth_buoyancy = nullptr, ***dzeta_advect = nullptr;
    double ***dqv_advect = nullptr, ***dqc_advect = nullptr, ***dqr_advect = nullptr, ***dqr_VT = nullptr;
    double *dth_advectcont = nullptr, *dth_buoyancycont = nullptr, *dzeta_advectcont = nullptr;
    double *dqv_advectcont = nullptr, *dqc_advectcont = nullptr, *dqr_advectcont = nullptr, *dqr_VTcont = nullptr;

    // per-step phase timers, TIMEROUTPUTSIZE long (src/Declare.hpp:695-699)
    double *t_advection = nullptr, *t_poisson = nullptr, *t_diffusion = nullptr, *t_microphysics = nullptr,
           *t_all = nullptr;

    // Device back-end.  `dev` owns the device copy of the state;
    // `host_stale` is set while the device holds steps the host arrays
    // have not seen, `dev_stale` while host edits have not been uploaded.
    vvm_dev_t *dev = nullptr;
    vvm_dev_config dev_config{};
    int gpu_id = 0;
    bool host_stale = false;
    bool dev_stale = true;
    bool print_steps = true;           // `std::cout << step` each step, as the reference
    std::string sounding_path;         // LOADINITPATH of the reference; empty: analytic sounding
    std::string restart_prev_nc, restart_now_nc;  // LOADPATH1 / LOADPATH2 of the reference
    int start_step = 0;                // TIMENOW of the reference (TimeMarching's first step)
    bool output_nc_files = true;       // TimeMarching writes nc/<step>.nc on output steps
    bool output_nc_same_file = false;  // OUTPUTNCSAMEFILE: append [t][x][z] slices to nc/<step/1200000>.nc
    bool output_microphysics = false;  // OUTPUTMICROPHYSICS: Kessler rates in the same-file output
    bool output_state_files = false;   // ... and state/state_<step>.bin (exact restart)
    // Radiation coupling (src/Iteration.cpp:427-431, src/Rrtmgp.cpp): with a
    // solve_radiation set, TimeMarching calls it after every step that is a
    // multiple of radiation_every, with th, qv, qc, qr and zeta holding the
    // values the reference's Radiation::solve_radiation reads there (the
    // step's current time level, before nextTimeStep), and passes the
    // radiation_heating_rate it leaves to the device for the next steps.
    // radiation_cooling is RADIATIONCOOLING (-2 K/day on thp).
    std::function<void(vvm &)> solve_radiation;
    int radiation_every = 40;
    bool radiation_cooling = false;
    // Column microphysics on the host in place of Kessler (the reference's
    // P3 call, src/Iteration.cpp:314-402): TimeMarching hands it each
    // step's fields in the [i][k] layout (see vvm_dev_set_host_microphysics)
    // and takes back thp, qvp, qcp, qrp and the process rates; it returns 0.
    std::function<int(vvm_host_view &)> host_microphysics;
    long long solve_iters_total = 0;   // PCG iterations over the run
    vvm_host_view host_view();          // pointers into this model's arrays
    void sync_to_device();              // upload when dev_stale
    void sync_to_host();                // download when host_stale

    // Host-side halo rules of src/Boundary.cpp:3-33 (periodic x; centre
    // rows copied, west/down rows zeroed), for arrays the caller edits.
    static void BoundaryProcess1D_center(double *var, int nz);
    static void BoundaryProcess2D_center(double **var, int nx, int nz);
    static void BoundaryProcess2D_westdown(double **var, int nx, int nz);
    // src/Boundary.cpp:34-52, on the device state.
    static void BoundaryProcess2D_all(vvm &model);

    // The forcing of src/AddForcing.cpp:6-28 (Q1LS/Q2LS when tropical_forcing
    // is set, the radiation heating rate, the random theta kick while
    // status_for_adding_forcing); the caller applies the halos, as
    // src/Iteration.cpp:283-285 does.
    static void AddForcing(vvm &model);

    // The reference's operators one at a time (src/Declare.hpp:725-749), each
    // on the device with exactly the reference function's writes.  Array
    // arguments must be the model's own arrays (the device holds their
    // copies); any other array is an error.
    static void Advection_zeta(vvm &model);                      // Advection.cpp:84-149
    static void Advection_thermo(double **past, double **now, double **future, double ***dvar,
                                 vvm &model);                     // Advection.cpp:25-81
    static void Advection_qrVT(vvm &model);                      // Advection.cpp:153-192
    static void Bouyancy(vvm &model);                            // Buoyancy.cpp:11-58

    class PoissonSolver {
       public:
        // Creates the device back-end (matrix-free: the operator is held as
        // its 1-D coefficient profiles) and uploads the state.
        static void InitPoissonMatrix(vvm &model);
        static void cal_w(vvm &model, int p = 0, int i = 0, int j = 0);
        static void cal_u(vvm &model);
        static void pubarTop_pt(vvm &model);
    };

    class Turbulence {
       public:
        // K_M, K_H (src/Turbulent.cpp:5-66) and the in-place diffusion of
        // zeta, th, qv, qc, qr (Mparam / Hparam, src/Turbulent.cpp:68-114).
        static void RKM_RKH(vvm &model);
        // The in-place diffusion of one field (var_future: zetap for Mparam,
        // thp, qvp, qcp or qrp for Hparam; var_now is unused, as in the
        // reference), src/Turbulent.cpp:68-114.
        static void Mparam(vvm &model, double **var_now, double **var_future);
        static void Hparam(vvm &model, double **var_now, double **var_future);
    };

    class NumericalProcess {
       public:
        // Nudge_theta, Nudge_zeta (unless CASE == 2) and Nudge_qv
        // (src/NumericalProcess.cpp:60-96, src/Iteration.cpp:415-417).
        static void Nudge(vvm &model);
        static void Nudge_theta(vvm &model);  // src/NumericalProcess.cpp:60-71
        static void Nudge_zeta(vvm &model);   // src/NumericalProcess.cpp:73-85
        static void Nudge_qv(vvm &model);     // src/NumericalProcess.cpp:87-96
    };

    class MicroPhysics {
       public:
        // NegativeValueProcess on qvp/qcp/qrp, condensation, autoconversion,
        // accretion, evaporation and the halos (src/Iteration.cpp:302-312,404)
        // in one device pass.
        static void Kessler(vvm &model);
        static void condensation(vvm &model);    // src/MicroPhysics.cpp:7-26
        static void autoconversion(vvm &model);  // src/MicroPhysics.cpp:29-51
        static void accretion(vvm &model);       // src/MicroPhysics.cpp:54-77
        static void evaporation(vvm &model);     // src/MicroPhysics.cpp:80-114
        // src/MicroPhysics.cpp:116-138 on one of the fifteen prognostic
        // arrays of a live model (found by its row pointers).
        static void NegativeValueProcess(double **var, int nx, int nz);
    };

    class Init {
       public:
        // src/Init.cpp:14-135: the analytic sounding, or the sounding file
        // model.sounding_path when set (LoadFile, src/Init.cpp:302-370).
        static void Init1d(vvm &model);
        // src/Init.cpp:137-266: initial fields from the profiles
        // (TROPICALFORCING branch when the config's tropical_forcing is set).
        static void Init2d(vvm &model);
        // src/Init.cpp:372-485: restart from model.restart_prev_nc and
        // model.restart_now_nc (two consecutive output_nc files) and the
        // sounding file model.sounding_path; TimeMarching then starts at
        // model.start_step (the reference's TIMENOW).
        static void LoadFromPreviousFile(vvm &model);
        // Base state and initial fields of the idealised benchmark cases
        // (warm bubble over the exponential-theta sounding); see vvm_init_case.
        static void InitCase(vvm &model, const vvm_case_params &p);
        static void RandomPerturbation(vvm &model, int t, double min_range = -0.25, double max_range = 0.25,
                                       double standard_deviation = 1.);
    };

    class Output {
       public:
        // <outputpath>/nc/<n>.nc as the reference's per-step output_nc
        // (src/Output.cpp:49-104): netCDF classic, dimensions x = nx and
        // z = nz, th, zeta, u, w, qv, qc, qr and the four Kessler process
        // rates as double [x][z], precip [x], and ubarTop (the reference
        // writes ubarTopp under that name).  With output_nc_same_file it
        // is the OUTPUTNCSAMEFILE variant (src/Output.cpp:106-285): one
        // file nc/<n/1200000>.nc with an unlimited dimension t, a record
        // t = n dt, th, zeta, u, w, qv, qc, qr as [t][x][z], ubarTop [t],
        // precip [t][x] (and the Kessler rates with output_microphysics),
        // created on first use and appended to afterwards.
        static void output_nc(int n, vvm &model);
        // <outputpath>/timer/<n>.nc: the phase timers t_advection,
        // t_poisson, t_diffusion, t_microphysics, t_all (src/Output.cpp:288-318).
        static void output_time_nc(int n, vvm &model);
        // <outputpath> with nc/, timer/ and state/ (src/Output.cpp:330-336).
        static void create_all_directory(vvm &model);
        // The model state at `step` as a VVMSTATE record file
        // <outputpath>/state/state_<step>.bin.
        static void output_state(int step, vvm &model);
        // The same record file at an explicit path.
        static void write_state(const std::string &path, vvm &model);
        // The base state on stdout and <outputpath>/init.txt, in the
        // reference's columns (src/Output.cpp:8-37); init2.txt, which lists
        // the P3/RRTMGP level profiles this model does not carry, is not written.
        static void printInit(vvm &model);
        // mkdir -p of one directory (src/Output.cpp:329-338).
        static void create_directory(std::string path);
    };

    class Iteration {
       public:
        static void pzeta_pt(vvm &model);
        // pth_pt, pqv_pt, pqc_pt and pqr_pt (src/Iteration.cpp:36-40,90-124)
        // act on independent fields; pthermo_pt runs the four as one device pass.
        static void pthermo_pt(vvm &model);
        static void pth_pt(vvm &model);
        static void pqv_pt(vvm &model);
        static void pqc_pt(vvm &model);
        static void pqr_pt(vvm &model);
        static void updateMean(vvm &model);  // src/Iteration.cpp:42-86
        static void TimeMarching(vvm &model);
        static void nextTimeStep(vvm &model);
    };

   private:
    void allocateMemory();
    void deallocateMemory();
};
