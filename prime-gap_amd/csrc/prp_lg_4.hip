// Instantiation unit: radix-2^28 variants, 8 lanes per number, S = 20 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 20> registered_lg_4;
}
