/*
 * Forwarding header for callers written against WFA2-lib's source tree:
 * `#include "wavefront/wavefront_align.h"` (wavefront/wavefront_align.h:34
 * pulls in the aligner, attributes, penalties and CIGAR declarations) resolves
 * here when include/ is on the include path, and declares the same API from
 * wfa_mi355x.h.  Link with libwfa2_amd.so instead of libwfa.a.
 */
#ifndef WFA_MI355X_WAVEFRONT_ALIGN_H_
#define WFA_MI355X_WAVEFRONT_ALIGN_H_
#include "../wfa_mi355x.h"
#endif
