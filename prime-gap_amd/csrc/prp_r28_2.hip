// Instantiation unit: radix-2^28 one-lane variants with digits {28, 30, 32}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<28, 30, 32> registered_r28_2;
}
