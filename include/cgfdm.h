/*
 * cgfdm.h — C ABI of the MI355X-native CGFDM time loop.
 *
 * A drop-in for the time loop that the reference drives from run()/propagate():
 * one opaque domain handle owns every device buffer of one simulation (wavefield,
 * packed metric/medium, PML auxiliaries, stations, PGV) on one GPU and one HIP stream.
 * Host buffers belong to the caller; device state belongs to the handle.
 *
 * Conventions
 *  - Every entry point returns an int status (CGFDM_OK == 0, negative on error) and never
 *    throws; cgfdm_last_error() returns a thread-local description of the last failure.
 *  - Grid indices are the reference's global interior indices: X in [0,NX), Y in [0,NY),
 *    Z in [0,NZ) with Z = NZ-1 the free surface (reference station.json / params.json).
 *  - Padded arrays have HALO = 3 cells on each side: _nx_ = NX+6 etc., x fastest
 *    (reference XFAST layout, axis.h:64-66).
 *  - Calls on one handle must be serialised by the caller; different handles are independent.
 */
#ifndef CGFDM_H
#define CGFDM_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFDM_HALO 3

enum cgfdm_status {
    CGFDM_OK = 0,
    CGFDM_EINVAL = -1,   /* bad argument / inconsistent descriptor            */
    CGFDM_ENOMEM = -2,   /* device or host allocation failed (reference 10001) */
    CGFDM_EHIP = -3,     /* HIP runtime error                                  */
    CGFDM_ECFL = -4,     /* DT above the CFL limit (reference calc_CFL, 110)   */
    CGFDM_EPML = -5,     /* nPML larger than the grid (reference codes 110/130)*/
    CGFDM_ESTATE = -6,   /* call out of order (e.g. step before set_model)     */
    CGFDM_EIO = -7       /* file could not be read or written                  */
};

/* Storage precision of the wavefield, RK accumulator and packed metric/medium.
 * Arithmetic is always FP32. FP16/BF16 use the reference's nondimensional scaling
 * (macro.h:65-73): V' = Cv*V, T' = Cs*T, buoyancy' = 1/(rho*Cs/Cv*Crho). */
enum cgfdm_precision {
    CGFDM_FP32 = 1,
    CGFDM_FP16 = 2,
    CGFDM_BF16 = 3
};

/* Domain descriptor: the subset of the reference PARAMS/GRID that the time loop reads
 * (struct.h:298-470, getParams.cpp:12-252). */
typedef struct cgfdm_desc {
    int NX, NY, NZ;        /* interior grid points                                   */
    int nPML;              /* PML thickness in cells (reference nPML, default 12)    */
    double DH;             /* grid spacing (m)                                       */
    double DT;             /* time step (s)                                          */
    int NT;                /* number of time steps = int(TMAX/DT) (propagate.cpp:36)  */
    int precision;         /* enum cgfdm_precision                                   */
    int free_surface;      /* 1: traction-free top, no top PML (reference FREE_SURFACE) */
    int pml;               /* 1: auxiliary-variable PML on the boundary faces (PML)  */
    int use_terrain;       /* 1: terrain-following vertical stretch from a height map */
    double depth;          /* model depth (m) used by the stretch (params Depth(km)*1e3) */
    int centerX, centerY;  /* coordinate origin in grid indices (params centerX/Y)   */
    double Cv, Cs, Crho;   /* scale factors; 0 selects the defaults of the precision */
    int check_cfl;         /* 1: reject DT > 1.34*Hmin/Vpmax (calc_CFL.cpp:331-352)  */
} cgfdm_desc;

typedef struct cgfdm_domain cgfdm_domain;

/* Returns a description of the last failure on this thread ("" if none). */
const char* cgfdm_last_error(void);

/* Creates a domain and allocates its device state (zeroed wavefield).
 * Replaces allocWave/allocPML/allocPMLParameter/allocCJM/allocMat_rDZ
 * (wave_deriv.cpp:24-48, pml_deriv.cpp:41-72, init_pml_para.cpp:21-51, CJM.cpp:22-39,
 * contravariant.cpp:179-200) and the grid setup of init_grid.cpp:14-130 for one rank. */
int cgfdm_domain_create(const cgfdm_desc* desc, int device, cgfdm_domain** out);

/* Frees every device buffer of the domain (freeWave/freePML/freeCJM...). */
int cgfdm_domain_destroy(cgfdm_domain* h);

/* Device bytes the domain holds (for admission by free HBM). */
long long cgfdm_domain_bytes(const cgfdm_domain* h);

/* Bytes cgfdm_domain_create would allocate for this descriptor (no allocation). */
long long cgfdm_estimate_bytes(const cgfdm_desc* desc);

/* Uploads the model and builds the packed metric/Jacobian/medium array in FP64 on the GPU,
 * the free-surface rDZ matrices and the PML profiles.
 *  terrain : height (m) of the surface on the padded (NY+6)x(NX+6) grid, x fastest;
 *            read only when desc->use_terrain (reference coord.cpp:255-307, terrain.cpp:930-937).
 *  vs,vp,rho: medium on the interior grid, NZ x NY x NX, x fastest (m/s, m/s, kg/m^3).
 * Replaces constructCoord/constructCJM/init_pml_parameter (coord.cpp:312-402,
 * CJM.cpp:170-236, contravariant.cpp:71-467, init_pml_para.cpp:99-285). */
int cgfdm_set_model(cgfdm_domain* h, const double* terrain,
                    const float* vs, const float* vp, const float* rho);

/* Single explosive Ricker source at a global interior index with peak frequency fc
 * (reference useSingleSource: locateSource/loadPointSource, singleSource.cpp:1-177,
 * propagate.cpp:145-146). M0 = 1e16 N*m as in the reference. */
int cgfdm_set_source_point(cgfdm_domain* h, int X, int Y, int Z, float fc);

/* Multi-point moment-rate source (reference useMultiSource: init_MultiSource,
 * calculateMomentRate, addMomenteRate; multiSource.cpp:824-922, addMoment.cpp:11-339).
 *  xyz : npts x 3 global interior indices (duplicates allowed; they are summed).
 *  rate: nt x npts x 6 moment-tensor rates (xx,yy,zz,xy,xz,yz) * area * slip rate,
 *        i.e. the table after solveMomentRate/changeStorageOrder, before the
 *        mu*Cv/(J*DH^3) scaling, which this call applies on the device. */
int cgfdm_set_source_table(cgfdm_domain* h, const int* xyz, const float* rate,
                           long long npts, int nt, float dt);

/* Stations (reference readStationIndex/initStationIndex, station.cpp:83-296):
 * xyz = n x 3 global interior indices. Points outside the grid are rejected. */
int cgfdm_set_stations(cgfdm_domain* h, const int* xyz, int n);

/* Advances nsteps full time steps (source, 4 RK stages, station and PGV sampling);
 * the body of the loop at propagate.cpp:142-198. Asynchronous on the domain's stream. */
int cgfdm_step(cgfdm_domain* h, int nsteps);

/* Waits for all work queued on the domain. */
int cgfdm_sync(cgfdm_domain* h);

/* Current time-step counter (number of completed steps). */
int cgfdm_current_step(const cgfdm_domain* h);

/* Station seismograms, n x NT x 9 float (Vx,Vy,Vz,Txx,Tyy,Tzz,Txy,Txz,Tyz), physical
 * units; the layout of station_mpi_X_Y_Z.bin after its int32 header (station.cpp:382-405). */
int cgfdm_get_stations(cgfdm_domain* h, float* out);

/* Peak ground velocity on the free surface, NY x NX x 2 float (PGVh, PGV);
 * the layout of PGV_Z_mpi_X_Y_Z.bin (PGV.cpp:103-111). */
int cgfdm_get_pgv(cgfdm_domain* h, float* out);

/* Plane of one wavefield variable (0..8), physical units (data_io.cpp:86-231):
 *  axis 0: X = index, out NZ x NY;  axis 1: Y = index, out NZ x NX;
 *  axis 2: Z = index, out NY x NX. */
int cgfdm_get_slice(cgfdm_domain* h, int axis, int index, int var, float* out);

/* Raw state access for tests and restarts: the 9 wavefield variables on the interior
 * grid (9 x NZ x NY x NX float, x fastest) in storage units (scaled by Cv/Cs). */
int cgfdm_get_state(cgfdm_domain* h, float* out);
int cgfdm_set_state(cgfdm_domain* h, const float* in);

/* The packed metric/Jacobian/medium array on the padded grid, 13 x num float
 * (xi_x..zt_z, J, mu, lambda, buoyancy), decoded from storage. */
int cgfdm_get_cjm(cgfdm_domain* h, float* out);

/* Average duration (ms) of the fused stage kernel launches timed with HIP events on
 * the domain's stream since the last reset, and the number of launches timed. */
int cgfdm_stage_timing(cgfdm_domain* h, int reset, double* avg_ms, long long* count);

/* Enables/disables HIP-event timing of the stage kernels (off by default). */
int cgfdm_enable_timing(cgfdm_domain* h, int on);

#ifdef __cplusplus
}
#endif

#endif /* CGFDM_H */
