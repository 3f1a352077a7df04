/*
 * Plain-data types that cross the C ABI of the MI355X sky-model predictor.
 *
 * The layouts are byte-for-byte those of hyperdrive's FFI structs in double
 * precision (src/gpu/types.h:35-275 with FLOAT = double), so the reference's
 * bindgen output (src/gpu/types_double.rs) describes them unchanged. Only the
 * double-precision form exists: the product path computes in FP64.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shapelet basis table geometry (src/gpu/types.h:23-26, src/model/shapelets.rs:11-14). */
#define SM_SBF_C 5000.0
#define SM_SBF_L 10001
#define SM_SBF_N 101
#define SM_SBF_DX 0.01
/* Reference frequency of power-law flux densities on the device (src/gpu/model.h:9). */
#define SM_POWER_LAW_FD_REF_FREQ 150e6

typedef struct RADec { double ra, dec; } RADec;     /* radians */
typedef struct HADec { double ha, dec; } HADec;     /* radians */
typedef struct XYZ { double x, y, z; } XYZ;          /* metres, geodetic */
typedef struct UVW { double u, v, w; } UVW;          /* metres */
/* 2*pi*(l, m, n - 1) of a component. */
typedef struct LmnRime { double l, m, n; } LmnRime;
/* Major/minor axis and position angle, radians. */
typedef struct GaussianParams { double maj, min, pa; } GaussianParams;
/* One shapelet coefficient; 16 bytes with trailing padding. */
typedef struct ShapeletCoeff {
    double value;
    unsigned char n1;
    unsigned char n2;
} ShapeletCoeff;
typedef struct ShapeletUV { double u, v; } ShapeletUV; /* metres */
/* Ionospheric offset constants of the peel loop (src/gpu/types.h:280-287): the phase
 * offset per metre of u and v, and the gain. */
typedef struct IonoConsts { double alpha, beta, gain; } IonoConsts;

/* 2x2 complex matrices, [XX, XY, YX, YY] with (re, im) interleaved; X = east-west. */
typedef struct JonesF32 {
    float j00_re, j00_im, j01_re, j01_im, j10_re, j10_im, j11_re, j11_im;
} JonesF32;
typedef struct JonesF64 {
    double j00_re, j00_im, j01_re, j01_im, j10_re, j10_im, j11_re, j11_im;
} JonesF64;

/* Common modelling inputs; every pointer is device memory (src/gpu/types.h:172-182). */
typedef struct Addresses {
    const int num_freqs;
    const int num_vis;
    const int num_baselines;
    const double *d_freqs;
    const double *d_shapelet_basis_values;
    const int num_unique_beam_freqs;
    const int *d_tile_map;
    const int *d_freq_map;
    const int *d_tile_index_to_unflagged_tile_index_map;
} Addresses;

/* Point components split by flux type: power law (instrumental FD at 150 MHz + SI), curved
 * power law (+ curvature q) and list (FDs precomputed [freq][component]) (types.h:187-208). */
typedef struct Points {
    const int num_power_laws;
    const LmnRime *power_law_lmns;
    const JonesF64 *power_law_fds;
    const double *power_law_sis;

    const int num_curved_power_laws;
    const LmnRime *curved_power_law_lmns;
    const JonesF64 *curved_power_law_fds;
    const double *curved_power_law_sis;
    const double *curved_power_law_qs;

    const int num_lists;
    const LmnRime *list_lmns;
    const JonesF64 *list_fds;
} Points;

typedef struct Gaussians {
    const int num_power_laws;
    const LmnRime *power_law_lmns;
    const JonesF64 *power_law_fds;
    const double *power_law_sis;
    const GaussianParams *power_law_gps;

    const int num_curved_power_laws;
    const LmnRime *curved_power_law_lmns;
    const JonesF64 *curved_power_law_fds;
    const double *curved_power_law_sis;
    const double *curved_power_law_qs;
    const GaussianParams *curved_power_law_gps;

    const int num_lists;
    const LmnRime *list_lmns;
    const JonesF64 *list_fds;
    const GaussianParams *list_gps;
} Gaussians;

/* Shapelet UVs are [baseline][component]; coefficients are flattened per component with
 * their counts in *_num_shapelet_coeffs (types.h:242-275). */
typedef struct Shapelets {
    const int num_power_laws;
    const LmnRime *power_law_lmns;
    const JonesF64 *power_law_fds;
    const double *power_law_sis;
    const GaussianParams *power_law_gps;
    const ShapeletUV *power_law_shapelet_uvs;
    const ShapeletCoeff *power_law_shapelet_coeffs;
    const int *power_law_num_shapelet_coeffs;

    const int num_curved_power_laws;
    const LmnRime *curved_power_law_lmns;
    const JonesF64 *curved_power_law_fds;
    const double *curved_power_law_sis;
    const double *curved_power_law_qs;
    const GaussianParams *curved_power_law_gps;
    const ShapeletUV *curved_power_law_shapelet_uvs;
    const ShapeletCoeff *curved_power_law_shapelet_coeffs;
    const int *curved_power_law_num_shapelet_coeffs;

    const int num_lists;
    const LmnRime *list_lmns;
    const JonesF64 *list_fds;
    const GaussianParams *list_gps;
    const ShapeletUV *list_shapelet_uvs;
    const ShapeletCoeff *list_shapelet_coeffs;
    const int *list_num_shapelet_coeffs;
} Shapelets;

/* ----------------------------------------------------------------- source-list interchange
 * A sky-model component as the host hands it to sm_new / sm_update_with_a_source. This has
 * no counterpart struct in the reference: it flattens SourceComponent
 * (src/srclist/types/components/mod.rs:21-31) for a C caller. */
typedef struct SmFluxDensity {
    double freq; /* Hz */
    double i, q, u, v; /* Jy */
} SmFluxDensity;

enum { SM_POINT = 0, SM_GAUSSIAN = 1, SM_SHAPELET = 2 };
enum { SM_LIST = 0, SM_POWER_LAW = 1, SM_CURVED_POWER_LAW = 2 };

typedef struct SmComponent {
    double ra, dec;          /* radians */
    int32_t comp_type;       /* SM_POINT / SM_GAUSSIAN / SM_SHAPELET */
    int32_t flux_type;       /* SM_LIST / SM_POWER_LAW / SM_CURVED_POWER_LAW */
    double maj, min, pa;     /* radians (Gaussians and shapelets) */
    double si, q;            /* spectral index, curvature */
    const SmFluxDensity *fds; /* power laws: fds[0] is the reference FD; lists sorted by freq */
    int32_t num_fds;
    int32_t num_coeffs;
    const ShapeletCoeff *coeffs;
} SmComponent;

#ifdef __cplusplus
} /* extern "C" */
#endif
