// Coarse-level kernel and its launcher in float.
#include "coarse_launch.hpp"

template void amgs::Engine<float>::coarse(int, int, int, const cplx<float>*, size_t, cplx<float>*, size_t, int, const amgs::CArgs<float>*);
