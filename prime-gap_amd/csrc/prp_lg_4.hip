// Instantiation unit: radix-2^26 variants, 8 lanes per number, S = 11 and 16 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<8, 11, 16> registered_lg_4;
}
