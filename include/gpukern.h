/*
 * gpukern.h -- C ABI of libgpukern.so, the MI355X (gfx950) build.
 *
 * Drop-in for the reference's kernel library: every function that the Go
 * binding (src/go/sim/gpu.go:34-262) calls through cgo is declared here with
 * the reference's name, argument order and meaning.  Each declaration cites
 * the reference header line it replaces.
 *
 * Contract kept from the reference:
 *   - device arrays are float* (held by Go as uintptr); vector fields are
 *     planar [3][X][Y][Z] (Z fastest), N = cells per component;
 *   - every call returns after the GPU work is complete (gpu_safe.h:41);
 *   - no return codes: errors print to stderr and abort() (gpu_safe.h:36-57);
 *   - the caller owns every buffer; plans are freed with delete_gpuFFT3dPlan.
 *
 * Implementation: hand-written HIP kernels for CDNA4 (wave64), own FFT
 * (no cuFFT/hipFFT), no CUDA compatibility layer.
 */
#ifndef GPUKERN_H
#define GPUKERN_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialisation (gpu_init.h:27-34, gpu_conf.h:27) ---- */
void gpu_init(int threads, int options);           /* gpu_init.h:27 */
void gpu_set_device(int devid);                    /* gpu_init.h:34 */
int gpu_maxthreads(void);                          /* gpu_conf.h:27 */
void gpu_print_properties_stdout(void);            /* gpu_properties.h:79 */

/* ---- memory (gpu_mem.h) ---- */
float* new_gpu_array(int size);                    /* gpu_mem.h:30  (zero-initialised) */
void free_gpu_array(float* ptr);                   /* gpu_mem.h:34 */
unsigned long long int gpu_usedmem(void);          /* gpu_mem.h:37 */
int gpu_stride_float(void);                        /* gpu_mem.h:52 */
void gpu_override_stride(int nFloats);             /* gpu_mem.h:60 */
#define DIR_TO 1
#define DIR_ON 2
#define DIR_FROM 3
void memcpy_gpu_dir(float* source, float* dest, int nElements, int direction); /* gpu_mem.h:89 */
void memcpy_to_gpu(float* source, float* dest, int nElements);                 /* gpu_mem.h:95 */
void memcpy_from_gpu(float* source, float* dest, int nElements);               /* gpu_mem.h:104 */
void memcpy_on_gpu(float* source, float* dest, int nElements);                 /* gpu_mem.h:113 */
float gpu_array_get(float* dataptr, int index);                                /* gpu_mem.h:124 */
void gpu_array_set(float* dataptr, int index, float value);                    /* gpu_mem.h:127 */
void gpu_zero(float* data, int nElements);                                     /* gpu_mem.h:134 */

/* ---- linear algebra (gpu_linalg.h:26-47) ---- */
void gpu_add(float* a, float* b, int N);                                        /* gpu_linalg.h:26 */
void gpu_madd(float* a, float cnst, float* b, int N);                           /* gpu_linalg.h:29 */
void gpu_madd2(float* a, float* b, float* c, int N);                            /* gpu_linalg.h:32 */
void gpu_add_constant(float* a, float cnst, int N);                             /* gpu_linalg.h:35 */
void gpu_linear_combination(float* a, float* b, float weightA, float weightB, int N); /* gpu_linalg.h:38 */
/* `vectors` and `weights` are HOST arrays of device pointers / weights, as
 * gpu.go:83-90 passes them (the reference kernel dereferenced them on the device). */
void gpu_linear_combination_many(float* result, float** vectors, float* weights, int NVectors, int NElem); /* gpu_linalg.h:44 */
void gpu_scale_dot_product(float* result, float* vector1, float* vector2, float a, int N); /* gpu_linalg.h:47 */

/* ---- reductions (gpu_reduction.h:24-38) ---- */
#define REDUCE_ADD 1
#define REDUCE_MAX 2
#define REDUCE_MAXABS 3
#define REDUCE_MIN 4
#define REDUCE_SUMABS 5
/* devbuffer/hostbuffer hold `blocks` partial results; blocks*threads*2 >= N */
float gpu_reduce(int operation, float* input, float* devbuffer, float* hostbuffer, int blocks,
                 int threadsPerBlock, int N);  /* gpu_reduction.h:31 */

/* ---- normalisation (gpu_normalize.h:26-31) ---- */
void gpu_normalize_uniform(float* m, int N);             /* gpu_normalize.h:26 */
void gpu_normalize_map(float* m, float* map, int N);     /* gpu_normalize.h:31 */

/* ---- torque (gpu_torque.h:27) ---- */
/* h <- dt_gilb * (m x h - alpha m x (m x h)); alpha = alpha_mul * alpha_map[i] (map may be NULL) */
void gpu_deltaM(float* m, float* h, float alpha_mul, float* alpha_map, float dt_gilb, int N); /* gpu_torque.h:27 */
void gpu_spintorque_deltaM(float* m, float* h, float alpha, float beta, float epsillon, float* u,
                           float* jmap, float dt_gilb, int N0, int N1, int N2); /* gpu_spintorque.h:23 */

/* ---- local fields (gpu_local_contr3.h:33-39) ---- */
/* Hext, anisK, anisAxes are HOST arrays (gpu.go:129-137); anisType 0 none, 1 uniaxial, 2 cubic */
void gpu_add_local_fields(float* m, float* h, int N, float* Hext, float* hmap, int anisType, float* anisK,
                          float* anisAxes); /* gpu_local_contr3.h:33 */
void gpu_add_local_fields_H_and_phi(float* m, float* h, float* phi, int N, float* Hext, float* hmap,
                                    int anisType, float* anisK, float* anisAxes); /* gpu_local_contr3.h:39 */
void gpu_add_lin_anisotropy(float* hx, float* hy, float* hz, float* mx, float* my, float* mz, float* kxx,
                            float* kyy, float* kzz, float* kyz, float* kxz, float* kxy, int N); /* gpu_anis.h:27 */

/* ---- exchange stencil (gpu_exch.h:23-29); size/periodic/exchInConv/cellSize are HOST int[3]/float[3] ---- */
#define EXCH_6NGBR 6
#define EXCH_12NGBR 12
void gpu_add_exch(float* m, float* h, int* size, int* periodic, int* exchInConv, float* cellSize,
                  int type); /* gpu_exch.h:23 */

/* ---- semi-analytical step (gpu_anal.h:34) ---- */
void gpu_anal_fw_step(float dt, float alpha, int Ntot, float* m_in, float* m_out, float* h); /* gpu_anal.h:34 */

/* ---- k-space kernel multiplication (gpu_kernmul.h:37-42) ---- */
void gpu_kernelmul6(float* fftMx, float* fftMy, float* fftMz, float* fftKxx, float* fftKyy, float* fftKzz,
                    float* fftKyz, float* fftKxz, float* fftKxy, int nRealNumbers); /* gpu_kernmul.h:37 */
void gpu_kernelmul4(float* fftMx, float* fftMy, float* fftMz, float* fftKxx, float* fftKyy, float* fftKzz,
                    float* fftKyz, int nRealNumbers); /* gpu_kernmul.h:42 */

/* ---- zero padding (gpu_zeropad.h:25-31) ---- */
void gpu_copy_pad(float* source, float* dest, int S0, int S1, int S2, int D0, int D1, int D2);   /* gpu_zeropad.h:25 */
void gpu_copy_unpad(float* source, float* dest, int S0, int S1, int S2, int D0, int D1, int D2); /* gpu_zeropad.h:31 */

/* ---- padded real FFT (gpu_fft0.h:30-75) ---- */
/* Layout of the transform: real [size] in, half-complex [paddedSize0][paddedSize1][paddedSize2+2]
 * floats out (interleaved complex, Z halved), as cufftPlan3d/2d R2C in the reference.
 * Rows and planes that are known zeros because of the padding are never transformed. */
typedef struct gpuFFT3dPlan gpuFFT3dPlan; /* gpu_fft0.h:30-38 (opaque here) */
gpuFFT3dPlan* new_gpuFFT3dPlan_padded(int* size, int* paddedSize);        /* gpu_fft0.h:51 */
void delete_gpuFFT3dPlan(gpuFFT3dPlan* plan);                             /* gpu_fft0.h:56 */
void gpuFFT3dPlan_forward(gpuFFT3dPlan* plan, float* input, float* output); /* gpu_fft0.h:64 */
/* unnormalised inverse; `input` is used as scratch and overwritten, as in the reference */
void gpuFFT3dPlan_inverse(gpuFFT3dPlan* plan, float* input, float* output); /* gpu_fft0.h:75 */
int gpuFFT3dPlan_normalization(gpuFFT3dPlan* plan);                        /* gpu_fft0.cu:112 */
/* No counterpart in the reference.  gpuFFT3dPlan_forward for callers that
 * cannot allocate device memory (HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same transform and the same bits, but
 * the temporaries of non-power-of-two lengths (packed rows, Bluestein buffers)
 * live in a fixed 32 MiB device array per GPU instead of per-call allocations,
 * and the plan's tables come from new_gpuFFT3dPlan_padded.  Plans whose
 * temporaries exceed 4 Mi complex elements are rejected (abort with a message).
 * Every call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads) race. */
void gpuFFT3dPlan_forward_noalloc(gpuFFT3dPlan* plan, float* input, float* output);

/* ---- thermal noise (gpu_temperature.h:25) ---- */
void gpu_gaussian_noise(float* devData, float mean, float stddev, int n); /* gpu_temperature.h:25 */

#ifdef __cplusplus
}
#endif
#endif /* GPUKERN_H */
