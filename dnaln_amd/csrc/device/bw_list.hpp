// bw_list.hpp — band half-widths with a compiled packed int16 specialisation (W = 2*bw + 1).
// Any other band width runs the generic int32 kernel.
#pragma once
#define DN_PACKED_BW_LIST(X) X(5) X(8) X(16) X(21) X(24) X(32) X(64)
// (bw, P): band widths whose packed fill also has a form split over P lanes per pair
// (sw_split.hpp), chosen per batch.
#define DN_SPLIT_LIST(X) X(32, 2) X(64, 4)
