// Instantiation unit: radix-2^28 variants, 16 lanes (a DPP row) per number, S = 5 and 10 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<16, 5, 10> registered_lg_5;
}
