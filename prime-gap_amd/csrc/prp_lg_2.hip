// Instantiation unit: radix-2^26 variants, 8 lanes per number, S = 18 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 18> registered_lg_2;
}
