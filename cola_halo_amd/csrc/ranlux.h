// Host stream of the RANLUX generator of the initial conditions (GSL ranlxd1
// semantics), for draws that persist across calls (subsample.c's generator).
#pragma once

namespace cola {

struct RanluxStream;
RanluxStream* ranlux_new(unsigned long long seed);
double ranlux_next(RanluxStream* s);
void ranlux_delete(RanluxStream* s);

}  // namespace cola
