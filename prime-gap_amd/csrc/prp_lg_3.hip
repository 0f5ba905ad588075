// Instantiation unit: radix-2^26 variants, 8 lanes per number, S = 21 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 21> registered_lg_3;
}
