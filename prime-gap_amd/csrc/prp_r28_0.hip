// Instantiation unit: radix-2^28 one-lane variants with digits {2, 4, 6, 8, 10, 12, 14, 16, 18}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<2, 4, 6, 8, 10, 12, 14, 16, 18> registered_r28_0;
}
