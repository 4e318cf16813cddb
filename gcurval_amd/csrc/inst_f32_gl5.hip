// Kernel instantiations for float arithmetic with 5-point Gauss-Legendre.
#include "launch_impl.h"

template struct gcurval::rt::Launch<float, 5>;
