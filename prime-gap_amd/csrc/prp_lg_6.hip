// Instantiation unit: radix-2^28 variants, a quad (4 lanes) per number, S = 38 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLG<4, 38> registered_lg_6;
}
