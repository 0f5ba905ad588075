// Instantiation unit: strong Lucas, 8 lanes per number, S = 4, 6 and 10 digits per lane.
#include "prp_lg_impl.h"
namespace pgap {
static RegisterLucas<8, 4, 6, 10> registered_lucas_0;
}
