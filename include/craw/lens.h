/*
 * libcraw for MI355X: the lens correction record, as a type only.
 *
 * CRAWFrameParams carries a CRAWLensCorrection by value
 * (include/craw/pipeline.h:312 and include/craw/lens.h of cyanreg/libcraw),
 * so the type is kept, byte for byte, for the frame parameters to keep
 * their 264-byte layout. The lens warp itself is not part of this library:
 * craw_pipeline_create() refuses CRAW_PIPELINE_FLAG_LENS with ENOTSUP, and
 * a frame's `lens` is ignored.
 */
#ifndef CRAW_LENS_H
#define CRAW_LENS_H

#include <craw/utils.h>

typedef struct CRAWLensCorrection {
    float k[3][4];      /* per channel R, G, B: k0..k3 */
    float centre[2];    /* optical centre, normalised frame coordinates */
    bool valid;
    uint8_t reserved[64 - 3*4*4 - 2*4 - 1];
} CRAWLensCorrection;

#endif /* CRAW_LENS_H */
