// Limits shared by the host factorisation and the device triangular solves.
#pragma once

namespace efem {

// widest supernode: the forward triangular-solve item stages one supernode's
// right-hand sides (3 x width values) in LDS
constexpr int kMaxSupernodeWidth = 768;

// device layout of the factor blocks: an m-row column occupies
// blockLd(m) = m rounded up to kBlockAlign rows (zero padding rows)
constexpr int kBlockAlign = 32;
constexpr int blockLd(int m) { return (m + kBlockAlign - 1) / kBlockAlign * kBlockAlign; }

}  // namespace efem
