// Instantiation unit: strong Lucas, one number per wavefront (three 16-lane rows computing a step's
// products at once), S = 2, 3, 5, 8, 10 and 11 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLucasWave<2, 3, 5, 8, 10, 11> registered_lucas_2;
}
