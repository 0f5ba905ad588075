// Instantiation unit: quad-per-number radix-2^26 variants, S = 24 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<24> registered_lg_0;
}
