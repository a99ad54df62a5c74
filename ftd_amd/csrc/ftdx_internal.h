// ftdx_internal.h -- error helpers shared by the libftdx.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace ftdx {

// Record `msg` as this thread's last error and return `code`.
int fail(int code, const std::string& msg);
// Record a HIP error (name and text) with context `what`; returns FTDX_ERR_HIP.
int hip_fail(hipError_t e, const char* what);

}  // namespace ftdx

#include <stddef.h>
#include <stdint.h>

typedef struct ftdx_context ftdx_context;

namespace ftdx {

// Tile configuration for the XEngine / FTD paths: `forced` if >= 0, else a shape
// heuristic (small problems use 64 x 64 tiles).
int xengine_cfg(int forced, int N, int batch);
// Bytes of pre-cast operand rows for batch x N rows of K samples.
size_t rows_bytes(int N, int K, int batch);
// One GGP payload: convert to pre-cast rows, then HERK into `out` in layout `fmt`
// (OUT_UPPER_PACKED or OUT_HERM_FULL), FP32 or FP64.
hipError_t run_ggp_payload(const uint8_t* in, void* out, uint8_t* rows, int N, int K, int batch,
                           int fmt, bool fp64, int cfg, int remap, hipStream_t s);
// Pre-cast conj(B) rows of a context whose prepare_b ran with FP8 compute and this
// (N, K, batch); nullptr otherwise.
const uint8_t* prepared_fp8_rows(const ftdx_context* ctx, int N, int K, int batch);
// One FTD raw-order payload -> packed lower triangle FP32.
hipError_t run_ftd_payload(const uint8_t* raw, float* tri, uint8_t* rows, int n_ant, int n_chan,
                           int n_time, int n_time_inner, int n_pol, bool twos, int cfg, int remap,
                           hipStream_t s);

}  // namespace ftdx
