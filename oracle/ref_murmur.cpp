// C entry for the reference's MurmurHash3_x86_32, built from its own MurmurHash3.cpp
// into oracle/_ref (test infrastructure: the checker of the type-id hash).
#include "MurmurHash3.h"

extern "C" void ref_murmur3_x86_32(const void* key, int len, unsigned int seed, void* out)
{
    MurmurHash3_x86_32(key, len, seed, out);
}
