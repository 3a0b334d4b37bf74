// Limits shared by the host factorisation and the device triangular solves.
#pragma once

namespace efem {

// widest supernode: the forward triangular-solve item stages one supernode's
// right-hand sides (3 x width values) in LDS
constexpr int kMaxSupernodeWidth = 768;

}  // namespace efem
