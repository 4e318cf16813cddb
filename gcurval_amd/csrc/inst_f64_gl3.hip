// Kernel instantiations for double arithmetic with 3-point Gauss-Legendre.
#include "launch_impl.h"

template struct gcurval::rt::Launch<double, 3>;
