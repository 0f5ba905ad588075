// Instantiation unit: quad-per-number radix-2^26 variants, S = 41 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<41> registered_lg_2;
}
