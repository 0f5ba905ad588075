// Instantiation unit: radix-2^28 one-lane variants with digits {46, 48}.
#include "prp_r28_impl.h"
namespace pgap {
static RegisterR28<46, 48> registered_r28_6;
}
