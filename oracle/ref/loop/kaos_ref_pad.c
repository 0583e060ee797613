/* kaos_ref_pad.c -- TEST INFRASTRUCTURE ONLY.  Heap padding for the reference
 * step-loop harness.  The reference's 3-D velocity grid generator allocates
 * Vperp1GpT with 2N - 1 elements and then writes element 2N and reads element
 * 0 (_0_0_VelGridGenerator.f90:72-91, N = Vperp12NlinRange), one element
 * past each end of the allocation; with the image's glibc that write lands on
 * the next chunk's header and aborts the run.  Linked with
 * -Wl,--wrap=malloc,--wrap=calloc,--wrap=realloc, every allocation gets 64
 * spare bytes after it and the out-of-range element falls into them; no value
 * the reference uses comes from there (element 2N is written again from
 * element 0 before anything reads it, and neither is inside the array). */
#include <stddef.h>

void* __real_malloc(size_t n);
void* __real_calloc(size_t m, size_t n);
void* __real_realloc(void* p, size_t n);

void* __wrap_malloc(size_t n) { return __real_malloc(n + 64); }
void* __wrap_calloc(size_t m, size_t n) { return __real_calloc(1, m * n + 64); }
void* __wrap_realloc(void* p, size_t n) { return __real_realloc(p, n + 64); }
