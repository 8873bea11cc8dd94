// Internal entry points shared between translation units of libflub.so
// (not part of the public C-ABI in include/flub.h).
#pragma once
#include <stdint.h>

// divergence_tag + mask_with_tag + scale_by_3d(s) in one pass over packed tags.
int flub_divergence_fused_u8(const float *u, const float *v, const float *w, float *div,
                             const float *ur, const float *vr, const float *wr, const uint8_t *tag,
                             int nx, int ny, int nz, float s, void *stream);
// The fused divergence that also writes the solver's right-hand side
// r = (FLUID ? d * k_r : 0) in the colour-split layout (div may be NULL), and
// when set the same values into rcopy and zeros into zero0 and zero1 (split
// arrays of the grid's size); divergence_split_ok tells whether the shapes and
// pointers allow it.
bool divergence_split_ok(const float *u, const float *v, const float *div, const uint8_t *tag, int nz);
int divergence_fused_split(const float *u, const float *v, const float *w, float *div, const float *ur,
                           const float *vr, const float *wr, const uint8_t *tag, int nx, int ny, int nz, float s,
                           float *rsplit, float k_r, void *stream, float *rcopy = nullptr, float *zero0 = nullptr,
                           float *zero1 = nullptr);
// subtract_gradient_tag reading the solver's x and applying the reference's
// pressure scale chain (x*h)*(1/h) on the fly.
int flub_subtract_gradient_x_u8(float *uf, const float *u_rigid, const float *x, const uint8_t *tag,
                                int nx, int ny, int nz, int d, float h, void *stream);
// The same face update from an FP64 x (the fused projection's FP64 solution,
// reference layout; p / h of the scale chain is x itself), computed in FP64
// and rounded once.
int flub_subtract_gradient_x64_u8(float *uf, const float *u_rigid, const double *x, const uint8_t *tag,
                                  int nx, int ny, int nz, int d, void *stream);
// The three axes of that update in one pass from the FP64 x in the
// colour-split layout (nz % 4 == 0; u, v and their rigid velocities 16-byte
// aligned: subtract_gradient3_split_ok).
bool subtract_gradient3_split_ok(const float *u, const float *v, const float *ur, const float *vr, int nz);
int subtract_gradient3_split(float *u, float *v, float *w, const float *ur, const float *vr, const float *wr,
                             const double *xs, const uint8_t *tag, int nx, int ny, int nz, void *stream);
