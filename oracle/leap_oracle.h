/*
 * leap_oracle.h — CPU restatement of LEAP-Accelerate's per-direction
 * calibration path, in plain C++17 FP64 with no third-party libraries.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * calibrator.  Only the test-suite, the smoke check and the CPU-baseline leg
 * of the benchmark may load it.  The product (the HIP calibrator) never calls
 * into it, and nothing here is a fallback for the product.
 *
 * Each function names the reference file:line it restates.  All paths are
 * relative to src/icrar/leap-accelerate/ of ska-telescope/icrar-leap-accelerate.
 *
 * Layout conventions (shared with the product's C ABI so a single Python
 * ctypes structure describes both):
 *   - matrices are column-major (Eigen's default), A is K x N, Ad is N x K;
 *   - the visibility cube is MS-native [row][channel][pol] complex, the layout
 *     of the DATA column; the reference transposes it to [channel][row][pol]
 *     when reading (ms/utils.h:133-145) which changes no arithmetic;
 *   - avgData is [baseline][pol] complex128 (row-major).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-memory measurement set: the accessor contract of icrar::MeasurementSet
 * used by the calibration path (MetaData.cc:46-89, Integration.cc:54-55).
 * Field order is identical to leap_ms_view in include/leap_hip.h. */
typedef struct oracle_ms_view {
    uint32_t num_stations;    /* MeasurementSet::GetNumStations */
    uint32_t total_antennas;  /* MeasurementSet::GetTotalAntennas (ANTENNA nrow) */
    uint32_t num_baselines;   /* MeasurementSet::GetNumBaselines */
    uint32_t num_rows;        /* MeasurementSet::GetNumRows */
    uint32_t num_channels;    /* SPECTRAL_WINDOW NUM_CHAN[0] */
    uint32_t num_pols;        /* POLARIZATION NUM_CORR[0] */
    double freq_start_hz;     /* SPECTRAL_WINDOW REF_FREQUENCY[0] */
    double freq_inc_hz;       /* SPECTRAL_WINDOW CHAN_WIDTH[0][0] */
    double phase_centre_ra_rad;   /* FIELD PHASE_DIR[0] */
    double phase_centre_dec_rad;
    const int32_t* antenna1;  /* ANTENNA1, at least the first num_baselines rows */
    const int32_t* antenna2;  /* ANTENNA2, at least the first num_baselines rows */
    const double* time;       /* TIME, num_rows values */
    const uint8_t* flag;      /* FLAG[row][chan 0][pol 0] for rows [0, num_baselines) */
    const double* uvw;        /* UVW, num_rows x 3 */
    const void* vis;          /* DATA, [num_rows][num_channels][num_pols] complex */
    int32_t vis_dtype;        /* 0 = complex128, 1 = complex64 */
    int32_t vis_on_device;    /* must be 0 for the oracle */
} oracle_ms_view;

typedef struct oracle_metadata oracle_metadata;

/* Error text of the last failing call on this thread. */
const char* oracle_last_error(void);

/* PhaseMatrixFunction (algorithm/cpu/PhaseMatrixFunction.cc:35-93).
 * ref_ant < 0 means "none".  A_out must hold (n+1) x (max(a1,a2)+1) doubles
 * (column-major, leading dimension n+1), I_out n ints.  On return *K is
 * A.rows() after trimming (the constraint row included) and *N is A.cols();
 * the first *K rows of A_out (leading dimension n+1) are valid.
 * Returns 0, or -1 with oracle_last_error() set when the reference throws. */
int oracle_phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg,
                        int32_t n, int32_t ref_ant,
                        double* A_out, int32_t* I_out, int32_t* K, int32_t* N);

/* SVDPseudoInverse (math/cpu/matrix_invert.h:49-56): thin SVD, singular values
 * below max(m,n)*eps*sigma_max dropped.  A is m x n column-major, Ad_out n x m. */
int oracle_pinv(const double* A, int32_t m, int32_t n, double* Ad_out);

/* Singular values (descending) of the m x n column-major A, for tests. */
int oracle_singular_values(const double* A, int32_t m, int32_t n, double* s_out);

/* MetaData::GenerateDDMatrix (model/cpu/MetaData.cc:166-203); dd is row-major 3x3. */
void oracle_dd_matrix(double phase_centre_ra, double phase_centre_dec,
                      double ra, double dec, double* dd_out);

/* Constants::GetChannelWavelength (model/cpu/MetaData.h:95-98). */
double oracle_channel_wavelength(double freq_start_hz, double freq_inc_hz, int32_t channel);

/* MeasurementSet::GetFilteredBaselines (ms/MeasurementSet.cc:181-240) as 0/1 bytes. */
int oracle_filtered_baselines(const oracle_ms_view* ms, double min_baseline_threshold,
                              uint8_t* out);

/* MetaData::MetaData (model/cpu/MetaData.cc:42-137).  ref_ant < 0 = none. */
int oracle_metadata_create(const oracle_ms_view* ms, int32_t ref_ant,
                           double min_baseline_threshold, oracle_metadata** out);
void oracle_metadata_destroy(oracle_metadata* md);
/* dims: [K, N, K1, nbl, reference_antenna] */
void oracle_metadata_dims(const oracle_metadata* md, int32_t* dims);
/* The degeneracy tests of MetaData.cc:129-136 ((Ad A).isApprox(I, 0.001) and
 * the same for A1) evaluated on the dense matrices: 1 = the reference warns. */
void oracle_metadata_degenerate(const oracle_metadata* md, int32_t* ad, int32_t* ad1);
/* Any pointer may be NULL.  Sizes follow the dims. */
void oracle_metadata_copy(const oracle_metadata* md, double* A, int32_t* I, double* Ad,
                          double* A1, int32_t* I1, double* Ad1);

/* CpuLeapCalibrator::RotateVisibilities (algorithm/cpu/CpuLeapCalibrator.cc:181-217)
 * for the integration rows [row_start, row_start + n_rows) and one direction.
 * avg_out is [nbl][pol] complex128, interleaved re/im; it is zeroed first
 * (CpuLeapCalibrator.cc:137). */
int oracle_rotate_visibilities(const oracle_metadata* md, const oracle_ms_view* ms,
                               uint32_t row_start, uint32_t n_rows,
                               const double* direction, double* avg_out);

/* The inversion half of CpuLeapCalibrator::PhaseRotate (CpuLeapCalibrator.cc:159-178)
 * from a given avgData [nbl][pol] complex128; cal_out has N entries. */
int oracle_apply_inversion(const oracle_metadata* md, const double* avg, double* cal_out);

/* CpuLeapCalibrator::Calibrate (CpuLeapCalibrator.cc:61-145).
 * Returns the number of solutions in *n_solutions (call with cal_out = NULL to
 * size the outputs); cal_out is [solution][direction][N], epochs_out is
 * [solution][2] (start, end).  ref_ant < 0 = none. */
int oracle_calibrate(const oracle_ms_view* ms, const double* directions, int32_t n_directions,
                     int32_t slice_start, int32_t slice_interval, int32_t slice_end,
                     double min_baseline_threshold, int32_t ref_ant,
                     double* cal_out, double* epochs_out, int32_t* n_solutions,
                     int32_t* n_antennas);

/* Number of solutions Slice(start, interval, end).Evaluate(timesteps).GetSize()
 * (common/Slice.cc:36-63, Slice.h:47-54, Range.h:40-63); -1 when it throws. */
int oracle_slice_solutions(int32_t start, int32_t interval, int32_t end, int32_t timesteps);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Calibration-residual loss and its gradient (test infrastructure for the
 * product's adjoint).  One solution interval over all rows, threshold 0.
 *   uvw(row)  = uvw0(row) + M_t (dp[a2] - dp[a1]),  M_t = enu_to_uvw[t] (3x3 row-major)
 *   direction = directions[d] + offsets[d]           (ra, dec radians)
 *   avgData, cal per direction as CpuLeapCalibrator (pol 0 reaches cal)
 *   loss      = sum_d sum_{n < K-1} wrap((A cal_d)(n) - arg avgData_d(I(n)))^2
 * Gradients by the chain rule written out with explicit sums (no recurrences).
 * antenna_delta may be NULL (zero).  grad_offsets [D][2], grad_antenna [N_ant][3]
 * (N_ant = max antenna index + 1); either may be NULL. */
int oracle_residual_loss_grad(const oracle_ms_view* ms, const double* directions, const double* offsets,
                              int32_t n_directions, const double* antenna_delta, const double* enu_to_uvw,
                              int32_t ref_ant, double* loss, double* grad_offsets, double* grad_antenna);

#ifdef __cplusplus
}
#endif
