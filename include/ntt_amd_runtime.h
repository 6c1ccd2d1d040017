/*
 * ntt_amd_runtime.h — minimal device-runtime helpers of the ntt_amd library.
 *
 * No counterpart in the reference, whose main() calls the CUDA runtime
 * directly (src/v04_montgomery.cu:947-967, 972-991).  These wrap the HIP
 * runtime the library itself links, so a binding without a HIP runtime of its
 * own (Python ctypes, or a process where another runtime copy is loaded) can
 * allocate operands, copy them and time launches on the same runtime the
 * kernels use.  All return an ntt_status (include/ntt_amd.h).
 */
#ifndef NTT_AMD_RUNTIME_H
#define NTT_AMD_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    char name[64];
    char arch[64];          /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int compute_units;
    int clock_khz;
    int l2_bytes;
    size_t lds_per_block;
    size_t total_global_mem;
} ntt_device_info;

int ntt_rt_device_count(int* count);
int ntt_rt_set_device(int device);
int ntt_rt_device_info(int device, ntt_device_info* out);
int ntt_rt_malloc(void** ptr, size_t bytes);
int ntt_rt_free(void* ptr);
int ntt_rt_memcpy_htod(void* dst, const void* src, size_t bytes);
int ntt_rt_memcpy_dtoh(void* dst, const void* src, size_t bytes);
/* Asynchronous on `stream`. */
int ntt_rt_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int ntt_rt_memset(void* ptr, int value, size_t bytes);
int ntt_rt_synchronize(void);
int ntt_rt_mem_info(size_t* free_bytes, size_t* total_bytes);
int ntt_rt_stream_create(void** stream);
int ntt_rt_stream_destroy(void* stream);
int ntt_rt_stream_synchronize(void* stream);
int ntt_rt_event_create(void** event);
int ntt_rt_event_destroy(void* event);
int ntt_rt_event_record(void* event, void* stream);
int ntt_rt_event_synchronize(void* event);
int ntt_rt_event_elapsed_ms(float* ms, void* start, void* stop);
/* Fills `words` uint32 values at `dst` with uniform residues of q_l in [0, q_l)
 * per limb (limb l = words [l*N, (l+1)*N)), from a counter-based hash of
 * (seed, global word index); `moduli_host` holds the limbs' q.  Used to make
 * large synthetic operands on the device without a host copy. */
int ntt_rt_fill_uniform(uint32_t* dst, const uint32_t* moduli_dev, size_t limbs, int logN,
                        size_t count, uint64_t seed, void* stream);
/* Butterfly roof: butterflies per second the whole device sustains running
 * the transform kernels' own butterfly code on register-resident operands
 * (no memory or LDS traffic).  variant 0: Cooley-Tukey, canonical; 1: CT,
 * lazy (q < 2^30); 2: Gentleman-Sande, canonical; 3: GS, lazy; 4: CT and
 * 5: GS on 64-bit words (q < 2^62, the ntt_plan64 butterflies); 6: CT, lazy,
 * with a precomputed-quotient (Shoup) product and 7: the same loop with the
 * Montgomery product (a comparison of the two products; eight twiddles per
 * thread used in turn).  Blocks until done (about a millisecond). */
int ntt_rt_butterfly_roof(int variant, double* butterflies_per_s, void* stream);
/* Plain streaming copy of `bytes` (multiple of 16, 16-byte aligned): reads and
 * writes each byte once; the reference kernel for bandwidth and for scaling
 * the HBM byte counters. */
int ntt_rt_stream_copy(void* dst, const void* src, size_t bytes, void* stream);
/* hipGraph capture of any sequence of this library's asynchronous calls on
 * `stream` (launchers, plan runs, batches, copies): begin, enqueue the calls,
 * end to get a replayable graph, launch it any number of times, destroy it.
 * The stream must not be the legacy default stream (it cannot be captured).
 * Operand addresses are baked into the graph. */
int ntt_rt_graph_begin(void* stream);
int ntt_rt_graph_end(void* stream, void** graph);
int ntt_rt_graph_launch(void* graph, void* stream);
int ntt_rt_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif

#endif /* NTT_AMD_RUNTIME_H */
