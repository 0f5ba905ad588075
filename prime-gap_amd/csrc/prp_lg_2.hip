// Instantiation unit: radix-2^28 variants, 8 lanes per number, S = 10 and 16 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 10, 16> registered_lg_2;
}
