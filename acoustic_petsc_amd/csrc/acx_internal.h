// acx_internal.h -- host-side declarations shared by the library's
// translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "../../include/acx.h"

namespace acx {

// Arguments of the fused linear stage kernel (acx_linear.hip).
struct LinArgs {
    const void *q;     // stage input (F32C: the HI part of the base state)
    const void *d;     // F32C: increment added to q before reconstruction (may be null)
    const void *u0;    // COMB: base state
    const void *lo;    // F32C final stage: LO part of the base state
    void *out;         // result
    void *out2;        // F32C final stage: new LO
    int nx, ny, nz, zc;
    int ntx, nty, ntz;   // tile counts (filled by the launcher)
    int bc[3];
    int final_stage;
    double a, b, c;    // epilogue coefficients
    double K, invrho, s, inv48h, lapc;
    unsigned *nf;      // non-finite watch word (host-mapped; null: no watch)
};

// Arguments of the fused nonlinear CWENO4 stage kernel (acx_cweno.hip).
struct CwArgs {
    const void *q;
    const void *d;
    const void *u0;
    const void *lo;
    void *out;
    void *out2;
    const uint8_t *flags;   // selective mode: per-cell bitmask, bit v set => full CWENO for component v
    const int *tlist;          // selective mode: marked tiles (count, then indices); null: every tile
    int nx, ny, nz, zc;
    int ntx, nty, ntz;
    int bc[3];
    int final_stage;
    double a, b, c;
    double K, invrho, s, inv_h, c2;
    double g4, g3c, g3s, inv_g4;   // linear weights: gamma_R4, gamma_R3[0], gamma_R3[1..6], 1/gamma_R4
    unsigned *nf;                  // non-finite watch word (host-mapped; null: no watch)
};

int launch_linear(const LinArgs &a, int sys, int epi, bool fp32, hipStream_t s);
// One whole SSP-RK3 (s = 4) step IN -> OUT of the linear ACOUSTIC4 operator on
// a periodic one-plane grid (acx_step2d.hip); IN and OUT distinct.
int launch_step2d(const LinArgs &a, const double *IN, double *OUT, double dt, hipStream_t s);
int launch_cweno(const CwArgs &a, int sys, int epi, bool fp32, bool selective, hipStream_t s);
// Tiles of the FP32 selective CWENO launch for a.zc: count, and the patch
// buffer for the flags: one byte per tile (1 for every tile holding a flagged
// cell or a face neighbour of one), then, from the next 4-byte boundary, the
// marked tiles as int32 (their number, then their indices in ascending order).
size_t cweno_sel_tiles(const CwArgs &a);
constexpr size_t sel_list_offset(size_t tiles) { return (tiles + 3) & ~(size_t)3; }
constexpr size_t sel_buffer_bytes(size_t tiles) { return sel_list_offset(tiles) + 4 * (tiles + 1); }
int launch_cweno_tilemask(const CwArgs &a, const uint8_t *flags, uint8_t *mask, hipStream_t s);

int set_error(int code, const char *fmt, ...);
// z-chunk of a 2.5D march from the round model (acx_api.hip); CUs of the current device
int auto_zc(int ntx, int nty, int nz, int slots, int start);
int device_cus();
// integer tuning knob from the environment (dflt when unset or empty)
int env_int(const char *name, int dflt);
int hip_check(hipError_t e, const char *what);
// Non-finite watch of the current device: the device address of its
// host-mapped flag word (nonfinite_word: created on first use, null if that
// fails; nonfinite_word_current: the existing word or null, never
// allocates), and whether a kernel has raised it (a host read, no
// synchronisation).
unsigned *nonfinite_word();
unsigned *nonfinite_word_current();
bool nonfinite_raised();

} // namespace acx
