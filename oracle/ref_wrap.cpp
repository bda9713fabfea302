// ref_wrap.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.c's header for who may use it).
//
// Compiles the reference's own OpenMP CPU closure from its source file where it lies
// (the path arrives as SE_REFERENCE_SOURCE from oracle/Makefile; nothing is copied) and
// exports plain C entry points so tests can cross-check the C restatement in oracle.c.
// SCALING_ELECTIONS_TEST makes the reference file skip its pybind11 module; the file's
// own test main() is compiled into the library but never called.
#include SE_REFERENCE_SOURCE

extern "C" {

// compute_strongest_paths_openmp<T, false>: scaling_elections.cu:738-849 (aligned variant;
// the tail variant reads uninitialised tile memory and is not exposed).  Requires
// n % tile == 0.  Returns 0, or -1 for an unsupported tile / misaligned size.
int ref_closure_openmp(const std::uint32_t* d, std::uint32_t n, std::uint32_t stride, std::uint32_t* p,
                       std::uint32_t tile) {
    if (tile == 0 || n % tile != 0) return -1;
    auto* prefs = const_cast<std::uint32_t*>(d);  // read only by the reference
    switch (tile) {
    case 4: compute_strongest_paths_openmp<4, false>(prefs, n, stride, p); break;
    case 8: compute_strongest_paths_openmp<8, false>(prefs, n, stride, p); break;
    case 16: compute_strongest_paths_openmp<16, false>(prefs, n, stride, p); break;
    case 32: compute_strongest_paths_openmp<32, false>(prefs, n, stride, p); break;
    case 64: compute_strongest_paths_openmp<64, false>(prefs, n, stride, p); break;
    default: return -1;
    }
    return 0;
}

}  // extern "C"
