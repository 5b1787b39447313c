// Host-side counting over batch buffers (the cut rules of the device parsers
// scan whole batches on the host before they are sent).
#pragma once

#include "common.hpp"

namespace sbwt_amd {

// Bytes equal to ch in p[0, n).
u64 count_byte(const char *p, u64 n, char ch);
// Little-endian u64 words equal to v among the n words at p (any alignment).
u64 count_word(const char *p, u64 n_words, u64 v);

}  // namespace sbwt_amd
