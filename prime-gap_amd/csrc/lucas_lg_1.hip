// Instantiation unit: strong Lucas, 16 lanes (a DPP row) per number, S = 5, 8, 10 and 11 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLucas<16, 5, 8, 10, 11> registered_lucas_1;
}
