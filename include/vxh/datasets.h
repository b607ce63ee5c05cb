/*
 * vxh datasets: real stereo pairs for the SGM path.
 *
 *   KITTI stereo  vx-vision/src/datasets/kitti.rs:11-107 loads one camera of
 *                 a KITTI odometry sequence (image_2, else image_0, :27-32)
 *                 with optional poses.  This loader keeps that behaviour
 *                 for the left camera and pairs every left frame with the
 *                 same file name in the matching right camera (image_3 with
 *                 image_2, image_1 with image_0), and reads the sequence's
 *                 calib.txt (projection matrices P0..P3) for the pair's
 *                 focal length, principal point and baseline.
 *   PNG           the frames' format: 8-bit gray, gray + alpha, RGB and RGBA,
 *                 non-interlaced (what KITTI ships), decoded with zlib.  The
 *                 luma conversion is that of the `image` crate 0.25
 *                 (DynamicImage::to_luma8, which the reference's examples use,
 *                 e.g. examples/feature_matching_demo.rs:43-48):
 *                 (2126 R + 7152 G + 722 B) / 10000, truncated; alpha is
 *                 dropped.
 */
#ifndef VXH_DATASETS_H
#define VXH_DATASETS_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vxh_kitti_stereo vxh_kitti_stereo;

/* One frame (DatasetFrame, vx-vision/src/datasets/mod.rs:12-24, for two
 * cameras).  The paths stay valid while the dataset lives. */
typedef struct {
    double timestamp;          /* frame index / 10 (KITTI records at ~10 Hz, kitti.rs:51) */
    const char *left_path;
    const char *right_path;
    int32_t has_pose;          /* 1 when the poses file has a line for this frame */
    float rotation[9];         /* R of the pose's 3x4 [R|t], row-major (kitti.rs:72-84) */
    float translation[3];      /* t of the pose */
} vxh_stereo_frame;

/* The pair's rectified camera from calib.txt: P0/P1 for the gray cameras,
 * P2/P3 for the color ones.  baseline = (P_left[3] - P_right[3]) / fx in the
 * units of the calibration (metres for KITTI). */
typedef struct {
    double fx, fy, cx, cy, baseline;
    int32_t valid;             /* 0 when calib.txt is absent or lacks the rows */
} vxh_stereo_calib;

/* KittiDataset::load (kitti.rs:19-68) for a stereo pair.  A missing or
 * unreadable image directory, or a left frame without its right image ->
 * VXH_ERR_INVALID_CONFIG (message names the path).  An unreadable poses
 * file gives frames without poses, as in the reference. */
vxh_status vxh_kitti_stereo_load(const char *sequence_dir, const char *poses_file,
                                 vxh_kitti_stereo **out);
void vxh_kitti_stereo_destroy(vxh_kitti_stereo *ds);
/* DatasetIterator::len (kitti.rs:103-107). */
size_t vxh_kitti_stereo_len(const vxh_kitti_stereo *ds);
/* 1 for the color cameras (image_2 / image_3), 0 for gray. */
int32_t vxh_kitti_stereo_is_color(const vxh_kitti_stereo *ds);
/* Frame i in file-name order (the iterator's i-th item, kitti.rs:90-100);
 * i >= len -> VXH_ERR_INVALID_CONFIG. */
vxh_status vxh_kitti_stereo_frame(const vxh_kitti_stereo *ds, size_t i, vxh_stereo_frame *out);
vxh_status vxh_kitti_stereo_calib(const vxh_kitti_stereo *ds, vxh_stereo_calib *out);
/* Decodes frame i's two images as luma (w*h bytes each, len = w*h; both
 * images must have the frame's size): the input the SGM path takes. */
vxh_status vxh_kitti_stereo_read_pair(const vxh_kitti_stereo *ds, size_t i, uint8_t *left,
                                      uint8_t *right, size_t len);

/* PNG size and channel count (1 gray, 2 gray + alpha, 3 RGB, 4 RGBA). */
vxh_status vxh_png_info(const char *path, uint32_t *w, uint32_t *h, uint32_t *channels);
/* Decoded samples, channels interleaved, rows top to bottom (len = w*h*channels). */
vxh_status vxh_png_read(const char *path, uint8_t *dst, size_t len);
/* DynamicImage::to_luma8 of the file (len = w*h). */
vxh_status vxh_png_read_luma8(const char *path, uint8_t *dst, size_t len);

#ifdef __cplusplus
}
#endif

#endif
