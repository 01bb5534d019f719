// synth.cpp -- seeded synthetic I420 clips (include/pcme_synth.h).
//
// Per clip, from splitmix64(seed):
//   * global pan (px, py) ~ U[-pan_max, pan_max] px/frame;
//   * luma background: U[0,1) noise on a margin-extended canvas, separable
//     Gaussian blur sigma = 3 (radius 9, edge-clamped), min/max scaled to
//     [40, 215]; frame t samples it bilinearly at
//     (c + (x - c) / z^t + pan * t) with z = 1 + zoom;
//   * nrect rectangles, sides ~ U{lo..hi}, flat value ~ U{0..255}, top-left
//     ~ U over the frame, velocity ~ U[-vmax, vmax]^2, drawn in order at
//     round(p0 + v t);
//   * N(0, noise_sigma) per luma pixel (Box-Muller on a per-frame stream),
//     rounded and clipped to [0, 255];
//   * chroma: an independent blurred-noise canvas at half resolution,
//     scaled to [108, 148], panned at half speed.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/pcme.h"
#include "../../include/pcme_synth.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * u01(); }
  int uint_in(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
};

struct Canvas {
  int w = 0, h = 0;
  std::vector<float> px;
  float at(int x, int y) const {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return px[(size_t)y * w + x];
  }
  float bilinear(double x, double y) const {
    double fx = floor(x), fy = floor(y);
    int ix = (int)fx, iy = (int)fy;
    float ax = (float)(x - fx), ay = (float)(y - fy);
    float a = at(ix, iy), b = at(ix + 1, iy), c = at(ix, iy + 1), d = at(ix + 1, iy + 1);
    return (a * (1 - ax) + b * ax) * (1 - ay) + (c * (1 - ax) + d * ax) * ay;
  }
};

// Blurred noise canvas scaled to [lo, hi].
static Canvas blurred_noise(Rng &rng, int w, int h, float lo, float hi) {
  Canvas c;
  c.w = w;
  c.h = h;
  c.px.resize((size_t)w * h);
  for (auto &v : c.px) v = (float)rng.u01();
  const int R = 9;
  const double sigma = 3.0;
  float k[2 * R + 1];
  double sum = 0;
  for (int i = -R; i <= R; i++) sum += exp(-(double)(i * i) / (2 * sigma * sigma));
  for (int i = -R; i <= R; i++) k[i + R] = (float)(exp(-(double)(i * i) / (2 * sigma * sigma)) / sum);
  std::vector<float> tmp((size_t)w * h);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float s = 0;
      for (int i = -R; i <= R; i++) s += k[i + R] * c.at(x + i, y);
      tmp[(size_t)y * w + x] = s;
    }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float s = 0;
      for (int i = -R; i <= R; i++) {
        int yy = std::min(std::max(y + i, 0), h - 1);
        s += k[i + R] * tmp[(size_t)yy * w + x];
      }
      c.px[(size_t)y * w + x] = s;
    }
  float mn = c.px[0], mx = c.px[0];
  for (float v : c.px) { mn = std::min(mn, v); mx = std::max(mx, v); }
  float sc = mx > mn ? (hi - lo) / (mx - mn) : 0.f;
  for (auto &v : c.px) v = lo + (v - mn) * sc;
  return c;
}

struct Rect {
  int w, h, val;
  double x0, y0, vx, vy;
};

struct Clip {
  pcme_synth_cfg cfg;
  double panx, pany;
  int margin, cmargin;
  Canvas bg, cu, cv;
  std::vector<Rect> rects;
};

static void make_frame(const Clip &cl, int t, uint8_t *Y, uint8_t *U, uint8_t *V) {
  const int w = cl.cfg.w, h = cl.cfg.h;
  const double z = pow(1.0 + (double)cl.cfg.zoom, (double)t);
  const double cx = 0.5 * (w - 1), cy = 0.5 * (h - 1);
  std::vector<float> L((size_t)w * h);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      double sx = cx + (x - cx) / z + cl.panx * t + cl.margin;
      double sy = cy + (y - cy) / z + cl.pany * t + cl.margin;
      L[(size_t)y * w + x] = cl.bg.bilinear(sx, sy);
    }
  for (const Rect &r : cl.rects) {
    int rx = (int)lround(r.x0 + r.vx * t), ry = (int)lround(r.y0 + r.vy * t);
    int xa = std::max(rx, 0), xb = std::min(rx + r.w, w);
    int ya = std::max(ry, 0), yb = std::min(ry + r.h, h);
    for (int y = ya; y < yb; y++)
      for (int x = xa; x < xb; x++) L[(size_t)y * w + x] = (float)r.val;
  }
  Rng nr{cl.cfg.seed ^ (0xA5A5A5A5DEADBEEFull * (uint64_t)(t + 1))};
  const double sg = cl.cfg.noise_sigma;
  const size_t n = (size_t)w * h;
  for (size_t i = 0; i < n; i += 2) {
    double u1 = 1.0 - nr.u01(), u2 = nr.u01();
    double rad = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
    double g0 = rad * cos(th), g1 = rad * sin(th);
    for (int k = 0; k < 2 && i + k < n; k++) {
      double v = (double)L[i + k] + sg * (k ? g1 : g0);
      long q = lround(v);
      Y[i + k] = (uint8_t)(q < 0 ? 0 : (q > 255 ? 255 : q));
    }
  }
  const int cw = w / 2, ch = h / 2;
  const double ccx = 0.5 * (cw - 1), ccy = 0.5 * (ch - 1);
  for (int y = 0; y < ch; y++)
    for (int x = 0; x < cw; x++) {
      double sx = ccx + (x - ccx) / z + 0.5 * cl.panx * t + cl.cmargin;
      double sy = ccy + (y - ccy) / z + 0.5 * cl.pany * t + cl.cmargin;
      long qu = lround(cl.cu.bilinear(sx, sy)), qv = lround(cl.cv.bilinear(sx, sy));
      U[(size_t)y * cw + x] = (uint8_t)std::min(std::max(qu, 0L), 255L);
      V[(size_t)y * cw + x] = (uint8_t)std::min(std::max(qv, 0L), 255L);
    }
}

}  // namespace

extern "C" int pcme_synth_clip(const pcme_synth_cfg *cfg, uint8_t *y, uint8_t *u, uint8_t *v,
                               int32_t threads) {
  if (!cfg || !y || !u || !v || cfg->w < 2 || cfg->h < 2 || (cfg->w & 1) || (cfg->h & 1) ||
      cfg->frames < 1 || cfg->nrect < 0 || cfg->rect_lo < 1 || cfg->rect_hi < cfg->rect_lo)
    return PCME_EINVAL;
  Clip cl;
  cl.cfg = *cfg;
  Rng rng{cfg->seed};
  cl.panx = rng.uni(-cfg->pan_max, cfg->pan_max);
  cl.pany = rng.uni(-cfg->pan_max, cfg->pan_max);
  cl.margin = (int)ceil(fabs((double)cfg->pan_max) * cfg->frames) + 4;
  cl.cmargin = cl.margin / 2 + 2;
  cl.bg = blurred_noise(rng, cfg->w + 2 * cl.margin, cfg->h + 2 * cl.margin, 40.f, 215.f);
  cl.cu = blurred_noise(rng, cfg->w / 2 + 2 * cl.cmargin, cfg->h / 2 + 2 * cl.cmargin, 108.f, 148.f);
  cl.cv = blurred_noise(rng, cfg->w / 2 + 2 * cl.cmargin, cfg->h / 2 + 2 * cl.cmargin, 108.f, 148.f);
  for (int i = 0; i < cfg->nrect; i++) {
    Rect r;
    r.w = rng.uint_in(cfg->rect_lo, cfg->rect_hi);
    r.h = rng.uint_in(cfg->rect_lo, cfg->rect_hi);
    r.val = rng.uint_in(0, 255);
    r.x0 = rng.uni(-0.5 * r.w, cfg->w - 0.5 * r.w);
    r.y0 = rng.uni(-0.5 * r.h, cfg->h - 0.5 * r.h);
    r.vx = rng.uni(-cfg->vmax, cfg->vmax);
    r.vy = rng.uni(-cfg->vmax, cfg->vmax);
    cl.rects.push_back(r);
  }
  int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, cfg->frames));
  const size_t ysz = (size_t)cfg->w * cfg->h, csz = ysz / 4;
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      int t = next.fetch_add(1);
      if (t >= cfg->frames) break;
      make_frame(cl, t, y + (size_t)t * ysz, u + (size_t)t * csz, v + (size_t)t * csz);
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nt; i++) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  return PCME_OK;
}
