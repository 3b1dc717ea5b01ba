// Order in which the camera-ray refill walks a task's (frame, pixel) items.
// Only the order is chosen here: each sample is seeded by its own frame and
// pixel (renderer.frag:129-142, rng.glsl:23-29), so any order that visits
// every (frame, pixel) of a task once gives the same samples.
#pragma once
#include "dev_math.hpp"

// A task's items over (frame of the task, region pixel): the first frame on
// its own, pixel by pixel (so a job's first frame completes as soon as it
// can), then the other frames in groups of kFrameGroup, each group walked
// chunk by chunk of 64 pixels (one traversal wave, an 8 x 8 tile of the
// tiled pixel order below) and, within a chunk, frame by frame: the waves
// that follow each other trace the same pixels with other jitter and find
// the same nodes in the caches. The mapping depends on the task's frame
// count, so a task's frame count must not change once any of its items was
// issued.
constexpr int kFrameGroup = 16;
DI void item_frame_pixel(int64_t local, int npix, int nframes, int& fslot, int& rpix) {
  if (local < npix || nframes <= 1) {
    fslot = (int)(local / npix);
    rpix = (int)(local - (int64_t)fslot * npix);
    return;
  }
  local -= npix;
  const int64_t gsz = (int64_t)kFrameGroup * npix;
  const int grp = (int)(local / gsz);
  const int gframes = min(kFrameGroup, nframes - 1 - grp * kFrameGroup);
  const int64_t l = local - (int64_t)grp * gsz;
  const int C = npix / 64, R = npix - C * 64;
  int f;
  if (l < (int64_t)C * 64 * gframes) {
    const int c = (int)(l / (64 * gframes));
    const int rem = (int)(l - (int64_t)c * 64 * gframes);
    f = rem / 64;
    rpix = c * 64 + rem % 64;
  } else {  // the last, partial chunk (R > 0 here)
    const int l2 = (int)(l - (int64_t)C * 64 * gframes);
    f = l2 / R;
    rpix = C * 64 + l2 % R;
  }
  fslot = 1 + grp * kFrameGroup + f;
}

// Region pixel -> offset (dx, dy) inside a render region of rw x rh pixels:
// 8 x 8 tiles in bands of 8 rows, tiles left to right, the last band and the
// last tile of a band cut to the region, so the 64 camera rays of a
// traversal wave start as a square of pixels.
DI void region_tile_pixel(int rpix, int rw, int rh, int& dx, int& dy) {
  const int band = rpix / (8 * rw), idx = rpix - band * 8 * rw;
  const int bh = min(8, rh - band * 8);
  const int t = idx / (8 * bh), j = idx - t * 8 * bh;
  const int tw = min(8, rw - t * 8);
  dx = t * 8 + j % tw;
  dy = band * 8 + j / tw;
}
