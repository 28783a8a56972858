/*
 * lzr_netinit.h -- the network's prediction with the net-start fallback's start-point
 * residual, shared by the network (lzr_net.hip) and the net-initialized fit (lzr_fit.hip).
 */
#ifndef LZR_NETINIT_H
#define LZR_NETINIT_H

#include "../../include/localizer_hip.h"

/* lzr_net_predict_device that also writes, per candidate, |f|^2 at the reference start point
 * (amplitude = intensity, background, width w_start, position = the candidate's) over the
 * network's 11 x 11 patch (x_offset = floor(x - 5), y likewise), NaN where the patch leaves
 * the frame, into d_start_sse (ncandidates doubles; NULL: not written).  The patch is the fit
 * window of a candidate whose half-width is 5, so the net-start fallback needs no second
 * read of those windows.  With d_nsel (a device count) only the candidates d_sel lists, in
 * its first *d_nsel entries, are predicted (their rows in list order); the others' starts and
 * residuals are left as they are. */
int lzr_net_predict_start_sse(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                              const LzrCandidate *d_cands, int ncands, LzrStart *d_starts, double w_start,
                              double *d_start_sse, const int *d_sel, const unsigned *d_nsel, void *stream);

#endif /* LZR_NETINIT_H */
