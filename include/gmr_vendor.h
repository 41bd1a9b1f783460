/* gmr_vendor.h -- C ABI of the vendor baselines on MI355X: the uncompressed
 * matrix in CSR form multiplied by rocSPARSE, and the GPU energy counter.
 *
 * These replace the reference's cuSPARSE baseline and NVML energy reads
 * (gpu-engine/cusparse_test.cu:264-502; energy at cusparse_test.cu:392-425 and
 * gpu_engine.cu:286-330). They are measurement tools beside the grammar
 * engine, not part of its sweep: nothing in libgmr calls them.
 *
 * Every function returns GMR_OK (0) or a gmr_status code; gmr_last_error()
 * (gmr_host.h) gives the message.
 */
#ifndef GMR_VENDOR_H
#define GMR_VENDOR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmr_csr gmr_csr;

enum gmr_csr_alg {
    GMR_CSR_ALG_DEFAULT = 0,   /* rocSPARSE's default for the format */
    GMR_CSR_ALG_ROW = 1,       /* SpMV: adaptive; SpMM: csr (row split, LDS) */
    GMR_CSR_ALG_ROW_SPLIT = 2, /* SpMV: row split; SpMM: csr_row_split (shuffles) */
    GMR_CSR_ALG_NNZ_SPLIT = 3  /* SpMM only: csr_merge (nnz split); GMR_ERR_INVALID for B = 1.
                                  An algorithm rocSPARSE rejects returns GMR_ERR_DEVICE */
};

/* Upload a CSR matrix (host arrays: row_ptr[rows+1], col_idx[nnz], vals[nnz],
 * 32-bit indices) to the current device. */
int gmr_csr_create(int rows, int cols, long long nnz, const int* row_ptr, const int* col_idx,
                   const float* vals, gmr_csr** out);
/* Y = M X on device pointers: X is cols x B and Y rows x B, row-major
 * (vectors innermost, as the engine's X[c*B+b]); B = 1 runs SpMV, B > 1 SpMM.
 * stream is a hipStream_t (NULL: the default stream). */
int gmr_csr_run(gmr_csr* m, const float* d_X, int B, float* d_Y, int alg, void* stream);
/* Host in / host out for drivers: copies X in, one untimed product, then
 * `iters` products timed with HIP events (*ms = mean per product), copies Y out. */
int gmr_csr_bench(gmr_csr* m, const float* h_X, int B, float* h_Y, int alg, int iters, double* ms);
/* Energy of B-vector products: products run back to back between two updates
 * of the energy counter at least min_seconds apart (the reference's >= 1.5 s
 * NVML loop, cusparse_test.cu:392-420); *mj = millijoules per product, *mw =
 * mean power in milliwatts. Uses device X / Y buffers of its own.
 * GMR_ERR_DEVICE when the counter cannot be read or does not advance. */
int gmr_csr_energy(gmr_csr* m, int B, int alg, double min_seconds, double* mj, double* mw);
/* Device bytes held: CSR arrays plus the largest rocSPARSE workspace so far. */
long long gmr_csr_device_bytes(const gmr_csr* m);
void gmr_csr_free(gmr_csr* m);

/* Accumulated energy of GPU `device` in microjoules, from the amdsmi energy
 * counter (the NVML total-energy read's counterpart), and the time of the
 * counter's last update in ns (may be NULL). The counter advances in steps
 * (about once a second on MI355X), so energy over an interval is taken
 * between two updates. GMR_ERR_DEVICE when the counter cannot be read. */
int gmr_energy_uj(int device, double* uj, unsigned long long* timestamp_ns);

#ifdef __cplusplus
}
#endif

#endif /* GMR_VENDOR_H */
