// Instantiation unit: radix-2^28 variants, a quad (4 lanes) per number, S = 24 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<4, 24> registered_lg_0;
}
