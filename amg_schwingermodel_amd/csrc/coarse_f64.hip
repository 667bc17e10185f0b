// Coarse-level kernel and its launcher in double.
#include "coarse_launch.hpp"

template void amgs::Engine<double>::coarse(int, int, int, const cplx<double>*, size_t, cplx<double>*, size_t, int, const amgs::CArgs<double>*);
