// Instantiation unit: radix-2^28 one-lane variants with digits {34, 36}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<34, 36> registered_r28_3;
}
