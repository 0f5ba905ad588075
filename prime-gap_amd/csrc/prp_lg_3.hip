// Instantiation unit: radix-2^28 variants, 8 lanes per number, S = 19 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 19> registered_lg_3;
}
