// Instantiation unit: radix-2^28 one-lane variants with digits {20, 22, 24, 26}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<20, 22, 24, 26> registered_r28_1;
}
