/* batch.hip -- host side of the batched persistent execution
 * (mpgfx_batch_create / _run / _info / _destroy, include/mpgfx.h).
 *
 * Planning groups the tasks into classes by (precision, real/complex), builds
 * each class's phase-1 item list and its phase-2 chain groups (sorted longest
 * first, R chains per wavefront), sizes the product scratch for the batch, and
 * captures a single hipGraph: per class [reset counters -> product kernel] on
 * a forked stream, then, after the join, one fold kernel for the fold units of
 * every class (sorted longest first across classes).  One graph launch runs
 * the whole batch.  Counters are reset by small kernels inside the graph (no
 * memset nodes).
 *
 * Tuning: the mapping (values per product item, lanes per workgroup,
 * workgroups per CU) of each class's products (keyed by precision,
 * real/complex and length bucket) and of the fold kernel (keyed by the number
 * of units and the longest chain, as log2 buckets) is either the fixed default
 * or the tuner's pick: each candidate is timed on the batch's own data, and
 * the winners are cached in a small JSON file.
 */
#include "common.h"
#include "dispatch.h"
#include "batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

struct mpgfx_batch
{
  struct cls
  {
    int pb, cplx;
    const mpgfx::pb_ops *ops;
    std::vector<mpgfx::bt_task> tasks;
    std::vector<std::pair<int, int64_t>> axpy_like;     /* (task, values) for phase-1 items */
    mpgfx::bt_class dev;                                 /* device view */
    void *d_tasks = nullptr, *d_items = nullptr, *d_chains = nullptr, *d_gbase = nullptr;
    void *d_scr_l = nullptr, *d_scr_s = nullptr, *d_counters = nullptr;
    int64_t total_values = 0, total_chain_steps = 0, nchains = 0;
    std::vector<int64_t> gmax;                           /* longest chain of each fold group */
    mpgfx_mapping map1;
    int bucket1 = 0;
  };
  std::vector<cls> classes;
  /* phase 2: fold units of all classes, longest first, and the class table */
  std::vector<mpgfx::bt_unit> units;
  void *d_units = nullptr, *d_classes = nullptr, *d_counter = nullptr;
  mpgfx_mapping map2;
  int bucket_units = 0, bucket_len = 0;
  std::vector<hipStream_t> side;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t scratch_bytes = 0;
  int cus = 256;
};

namespace mpgfx {
namespace {

int hip_ok (const char *where, hipError_t e)
{
  if (e != hipSuccess) { set_error (where, e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

const mpgfx_mapping DEFAULT_P1 = {256, 256, 4, 0};
const mpgfx_mapping DEFAULT_P2 = {0, 64, 8, 0};

int valid_task (const mpgfx_task &t)
{
  if (t.kind < MPGFX_TASK_AXPY || t.kind > MPGFX_TASK_GEMV) return MPGFX_EINVAL;
  if (!find_ops (t.pb)) return MPGFX_EUNSUPPORTED;
  if (t.n < 0 || t.m < 0 || t.n > ((int64_t) 1 << 40) || t.m > ((int64_t) 1 << 31)) return MPGFX_EINVAL;
  if (t.kind == MPGFX_TASK_AXPY)
    {
      if (t.n > 0 && !(t.in0_limbs && t.in0_se && t.in1_limbs && t.in1_se && t.in2_limbs && t.in2_se &&
                       t.out_limbs && t.out_se))
        return MPGFX_EINVAL;
    }
  else
    {
      bool has_out = t.kind == MPGFX_TASK_DOT || t.m > 0;
      if (has_out && !(t.out_limbs && t.out_se)) return MPGFX_EINVAL;
      if (has_out && t.n > 0 && !(t.in0_limbs && t.in0_se && t.in1_limbs && t.in1_se)) return MPGFX_EINVAL;
    }
  return MPGFX_OK;
}

/* build phase-1 items for one class with the given chunk */
std::vector<bt_item> make_items (const mpgfx_batch::cls &c, int chunk)
{
  std::vector<bt_item> items;
  for (auto &tv : c.axpy_like)
    for (int64_t s = 0; s < tv.second; s += chunk)
      {
        bt_item it;
        it.task = tv.first;
        it.start = s;
        it.count = (int32_t) std::min<int64_t> (chunk, tv.second - s);
        items.push_back (it);
      }
  /* larger tasks first: long products feed the long chains */
  return items;
}

int upload_items (mpgfx_batch::cls &c, int chunk, hipStream_t s)
{
  std::vector<bt_item> items = make_items (c, chunk);
  if (c.d_items) { (void) hipFree (c.d_items); c.d_items = nullptr; }
  c.dev.nitems = (i64) items.size ();
  if (!items.empty ())
    {
      int rc = hip_ok ("batch items", hipMalloc (&c.d_items, items.size () * sizeof (bt_item)));
      if (rc) return rc;
      rc = hip_ok ("batch items copy", hipMemcpyAsync (c.d_items, items.data (), items.size () * sizeof (bt_item),
                                                       hipMemcpyHostToDevice, s));
      if (rc) return rc;
      rc = hip_ok ("batch items sync", hipStreamSynchronize (s));
      if (rc) return rc;
    }
  c.dev.items = (const bt_item *) c.d_items;
  return MPGFX_OK;
}

int grid_of (const mpgfx_mapping &m, int cus) { return std::max (1, m.wg_per_cu * cus); }

int launch_class (mpgfx_batch::cls &c, int cus, hipStream_t s)
{
  int rc = c.ops->batch_launch (0, c.cplx, c.dev, 1, 64, s);
  if (!rc && c.dev.nitems > 0) rc = c.ops->batch_launch (1, c.cplx, c.dev, grid_of (c.map1, cus), c.map1.block, s);
  return rc;
}

int launch_folds (mpgfx_batch *b, const mpgfx_mapping &mp, hipStream_t s)
{
  return batch_folds_launch ((const bt_class *) b->d_classes, (const bt_unit *) b->d_units, (i64) b->units.size (),
                             (unsigned *) b->d_counter, grid_of (mp, b->cus), mp.block, s);
}

/* ---- tuning cache (flat JSON object: "key": [chunk, block, wg_per_cu]) ---- */
std::map<std::string, mpgfx_mapping> read_cache (const char *path)
{
  std::map<std::string, mpgfx_mapping> m;
  if (!path) return m;
  FILE *f = fopen (path, "r");
  if (!f) return m;
  std::string txt;
  char buf[4096];
  size_t r;
  while ((r = fread (buf, 1, sizeof buf, f)) > 0) txt.append (buf, r);
  fclose (f);
  size_t pos = 0;
  while ((pos = txt.find ('"', pos)) != std::string::npos)
    {
      size_t e = txt.find ('"', pos + 1);
      if (e == std::string::npos) break;
      std::string key = txt.substr (pos + 1, e - pos - 1);
      mpgfx_mapping mm = {0, 0, 0, 0};
      if (sscanf (txt.c_str () + e + 1, " : [ %d , %d , %d ]", &mm.chunk, &mm.block, &mm.wg_per_cu) == 3 &&
          mm.block > 0 && mm.block <= 256 && mm.block % 64 == 0 && mm.wg_per_cu > 0 && mm.wg_per_cu <= 64 &&
          mm.chunk >= 0 && mm.chunk <= (1 << 20))
        m[key] = mm;
      pos = e + 1;
    }
  return m;
}

void write_cache (const char *path, const std::map<std::string, mpgfx_mapping> &m)
{
  if (!path) return;
  std::string tmp = std::string (path) + ".tmp";
  FILE *f = fopen (tmp.c_str (), "w");
  if (!f) return;
  fprintf (f, "{\n");
  size_t i = 0;
  for (auto &kv : m)
    fprintf (f, "  \"%s\": [%d, %d, %d]%s\n", kv.first.c_str (), kv.second.chunk, kv.second.block,
             kv.second.wg_per_cu, ++i < m.size () ? "," : "");
  fprintf (f, "}\n");
  fclose (f);
  rename (tmp.c_str (), path);
}

std::string key_of (int phase, const mpgfx_batch::cls &c, int bucket)
{
  char b[96];
  snprintf (b, sizeof b, "p%d_pb%d_c%d_L%d", phase, c.pb, c.cplx, bucket);
  return b;
}

float time_phase (mpgfx_batch::cls &c, int phase, const mpgfx_mapping &mp, int cus, hipStream_t s, int reps)
{
  hipEvent_t e0, e1;
  if (hipEventCreate (&e0) != hipSuccess) return -1.f;
  if (hipEventCreate (&e1) != hipSuccess) { (void) hipEventDestroy (e0); return -1.f; }
  float best = -1.f;
  for (int r = 0; r < reps; r++)
    {
      int rc = c.ops->batch_launch (0, c.cplx, c.dev, 1, 64, s);
      if (!rc) rc = hip_ok ("tune", hipEventRecord (e0, s));
      if (!rc) rc = c.ops->batch_launch (phase, c.cplx, c.dev, grid_of (mp, cus), mp.block, s);
      if (!rc) rc = hip_ok ("tune", hipEventRecord (e1, s));
      if (!rc) rc = hip_ok ("tune", hipEventSynchronize (e1));
      float ms = 0.f;
      if (!rc && hipEventElapsedTime (&ms, e0, e1) == hipSuccess)
        best = (best < 0.f || ms < best) ? ms : best;
      if (rc) { best = -1.f; break; }
    }
  (void) hipEventDestroy (e0);
  (void) hipEventDestroy (e1);
  return best;
}

int tune_class (mpgfx_batch::cls &c, int cus, hipStream_t s, std::map<std::string, mpgfx_mapping> &cache, bool *dirty)
{
  /* phase 1 */
  std::string k1 = key_of (1, c, c.bucket1);
  if (c.dev.nitems > 0 || !c.axpy_like.empty ())
    {
      auto it = cache.find (k1);
      if (it != cache.end ()) c.map1 = it->second;
      else
        {
          const int chunks[3] = {256, 1024, 4096}, blocks[2] = {256, 64}, wgs[3] = {2, 4, 8};
          float best = -1.f;
          mpgfx_mapping bm = DEFAULT_P1;
          for (int ch : chunks)
            {
              int rc = upload_items (c, ch, s);
              if (rc) return rc;
              for (int bl : blocks)
                for (int wg : wgs)
                  {
                    mpgfx_mapping mm = {ch, bl, wg, 0};
                    float ms = time_phase (c, 1, mm, cus, s, 2);
                    if (ms > 0.f && (best < 0.f || ms < best)) { best = ms; bm = mm; }
                  }
            }
          c.map1 = bm;
          cache[k1] = bm;
          *dirty = true;
        }
      int rc = upload_items (c, c.map1.chunk, s);
      if (rc) return rc;
    }
  return MPGFX_OK;
}

/* phase 2: the all-class fold kernel's mapping (needs every class's products
 * in the scratch: run phase 1 once first) */
int tune_folds (mpgfx_batch *b, hipStream_t s, std::map<std::string, mpgfx_mapping> &cache, bool *dirty)
{
  if (b->units.empty ()) return MPGFX_OK;
  char kb[96];
  snprintf (kb, sizeof kb, "p2_all_U%d_L%d", b->bucket_units, b->bucket_len);
  auto it = cache.find (kb);
  if (it != cache.end ()) { b->map2 = it->second; return MPGFX_OK; }
  for (auto &c : b->classes)
    {
      int rc = launch_class (c, b->cus, s);
      if (rc) return rc;
    }
  hipEvent_t e0, e1;
  if (hipEventCreate (&e0) != hipSuccess) return MPGFX_EHIP;
  if (hipEventCreate (&e1) != hipSuccess) { (void) hipEventDestroy (e0); return MPGFX_EHIP; }
  const int blocks[3] = {64, 128, 256}, wgs[4] = {1, 2, 4, 8};
  float best = -1.f;
  mpgfx_mapping bm = DEFAULT_P2;
  int rc = MPGFX_OK;
  for (int bl : blocks)
    for (int wg : wgs)
      {
        mpgfx_mapping mm = {0, bl, wg, 0};
        float ms = 0.f;
        rc = hip_ok ("tune", hipEventRecord (e0, s));
        if (!rc) rc = launch_folds (b, mm, s);
        if (!rc) rc = hip_ok ("tune", hipEventRecord (e1, s));
        if (!rc) rc = hip_ok ("tune", hipEventSynchronize (e1));
        if (!rc) rc = hip_ok ("tune", hipEventElapsedTime (&ms, e0, e1));
        if (rc) break;
        if (best < 0.f || ms < best) { best = ms; bm = mm; }
      }
  (void) hipEventDestroy (e0);
  (void) hipEventDestroy (e1);
  if (rc) return rc;
  b->map2 = bm;
  cache[kb] = bm;
  *dirty = true;
  return MPGFX_OK;
}

/* While the tuner runs the batch's own products (many times), AXPY outputs go
 * to one scratch buffer: an AXPY in place (out == in2, y = a*x + y) must still
 * read its original y in the first real run.  on = false restores the task
 * tables. */
int redirect_axpy_outputs (mpgfx_batch *b, bool on, void **scratch)
{
  if (on)
    {
      size_t lw = 0, sw = 0;
      for (auto &c : b->classes)
        for (auto &t : c.tasks)
          if (t.kind == MPGFX_TASK_AXPY)
            {
              const size_t parts = c.cplx ? 2 : 1;
              lw = std::max (lw, (size_t) t.n * parts * (size_t) ((c.pb + 63) / 64));
              sw = std::max (sw, (size_t) t.n * parts);
            }
      if (lw == 0) return MPGFX_OK;
      int rc = hip_ok ("batch tuning outputs", hipMalloc (scratch, (lw + sw) * 8));
      if (rc) return rc;
      for (auto &c : b->classes)
        {
          std::vector<bt_task> tt = c.tasks;
          for (auto &t : tt)
            if (t.kind == MPGFX_TASK_AXPY)
              {
                t.outl = (u64 *) *scratch;
                t.outs = (i64 *) ((u64 *) *scratch + lw);
              }
          rc = hip_ok ("batch tuning tasks", hipMemcpy (c.d_tasks, tt.data (), tt.size () * sizeof (bt_task),
                                                        hipMemcpyHostToDevice));
          if (rc) return rc;
        }
      return MPGFX_OK;
    }
  int rc = MPGFX_OK;
  for (auto &c : b->classes)
    if (!rc && !c.tasks.empty ())
      rc = hip_ok ("batch tasks restore", hipMemcpy (c.d_tasks, c.tasks.data (), c.tasks.size () * sizeof (bt_task),
                                                     hipMemcpyHostToDevice));
  if (*scratch) (void) hipFree (*scratch);
  *scratch = nullptr;
  return rc;
}

void free_class (mpgfx_batch::cls &c)
{
  void *ps[] = {c.d_tasks, c.d_items, c.d_chains, c.d_gbase, c.d_scr_l, c.d_scr_s, c.d_counters};
  for (void *p : ps)
    if (p) (void) hipFree (p);
}

int bucket_of (double mean)
{
  return mean > 1.0 ? (int) std::lround (std::log2 (mean)) : 0;
}

} // namespace
} // namespace mpgfx

using namespace mpgfx;

extern "C" int mpgfx_batch_destroy (mpgfx_batch *b)
{
  if (!b) return MPGFX_EINVAL;
  if (b->exec) (void) hipGraphExecDestroy (b->exec);
  if (b->graph) (void) hipGraphDestroy (b->graph);
  for (auto &c : b->classes) free_class (c);
  void *ps[] = {b->d_units, b->d_classes, b->d_counter};
  for (void *p : ps)
    if (p) (void) hipFree (p);
  for (auto s : b->side) (void) hipStreamDestroy (s);
  delete b;
  return MPGFX_OK;
}

extern "C" int mpgfx_batch_create (const mpgfx_task *tasks, int64_t ntasks, int tuning, const char *cache,
                                   void *stream, mpgfx_batch **out)
{
  if (!out || ntasks < 0 || (ntasks > 0 && !tasks) || tuning < 0 || tuning > 1) return MPGFX_EINVAL;
  *out = nullptr;
  for (int64_t i = 0; i < ntasks; i++)
    {
      int rc = valid_task (tasks[i]);
      if (rc) return rc;
    }
  hipStream_t s = (hipStream_t) stream;
  mpgfx_batch *b = new mpgfx_batch ();
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice (&dev) == hipSuccess && hipGetDeviceProperties (&prop, dev) == hipSuccess)
    b->cus = prop.multiProcessorCount;
  /* ---- classes ---- */
  std::map<std::pair<int, int>, int> idx;
  for (int64_t i = 0; i < ntasks; i++)
    {
      const mpgfx_task &t = tasks[i];
      auto key = std::make_pair (t.pb, t.cplx ? 1 : 0);
      if (!idx.count (key))
        {
          idx[key] = (int) b->classes.size ();
          mpgfx_batch::cls c;
          c.pb = t.pb; c.cplx = t.cplx ? 1 : 0; c.ops = find_ops (t.pb);
          c.map1 = DEFAULT_P1;
          std::memset (&c.dev, 0, sizeof c.dev);
          b->classes.push_back (c);
        }
      auto &c = b->classes[idx[key]];
      bt_task d;
      d.kind = t.kind; d.cplx = c.cplx; d.m = t.kind == MPGFX_TASK_GEMV ? t.m : 1; d.n = t.n;
      d.in0l = t.in0_limbs; d.in0s = t.in0_se; d.in1l = t.in1_limbs; d.in1s = t.in1_se;
      d.in2l = t.in2_limbs; d.in2s = t.in2_se; d.outl = t.out_limbs; d.outs = t.out_se; d.p0 = -1;
      c.tasks.push_back (d);
    }
  int rc = MPGFX_OK;
  for (auto &c : b->classes)
    {
      /* chains: one per dot, one per GEMV row; sorted longest first (stable, so
       * the rows of one GEMV stay consecutive) */
      std::vector<bt_chain> ch;
      for (int ti = 0; ti < (int) c.tasks.size (); ti++)
        {
          const bt_task &t = c.tasks[ti];
          if (t.kind == MPGFX_TASK_AXPY)
            {
              if (t.n > 0) c.axpy_like.push_back ({ti, t.n});
              c.total_values += t.n;
              continue;
            }
          for (int64_t r = 0; r < t.m; r++)
            ch.push_back ({ti, (int32_t) r, t.n});
          if (t.n > 0 && t.m > 0) c.axpy_like.push_back ({ti, t.m * t.n});
          c.total_values += t.m * t.n;
        }
      std::stable_sort (ch.begin (), ch.end (), [] (const bt_chain &a, const bt_chain &b2) { return a.len > b2.len; });
      for (int64_t p = (int64_t) ch.size () - 1; p >= 0; p--) c.tasks[ch[p].task].p0 = p;   /* first position */
      c.nchains = (int64_t) ch.size ();
      /* fold groups of R chains: a wavefront folds R chains at once, G = 64/R
       * lanes per number */
      const int64_t R = 64 / coop_lanes (c.pb);
      int64_t ngroups = (c.nchains + R - 1) / R;
      std::vector<int64_t> gbase (ngroups > 0 ? ngroups : 1, 0);
      int64_t acc = 0;
      c.gmax.assign ((size_t) ngroups, 0);
      for (int64_t g = 0; g < ngroups; g++)
        {
          gbase[g] = acc;
          c.gmax[g] = ch[g * R].len;                      /* longest chain of the group */
          acc += R * c.gmax[g];
          c.total_chain_steps += c.gmax[g];
        }
      while ((int64_t) ch.size () < ngroups * R) ch.push_back ({-1, 0, 0});
      const int64_t scr_n = acc;                          /* records per part */
      const int nl = (c.pb + 63) / 64, parts = c.cplx ? 2 : 1;
      /* product items: larger tasks first */
      std::stable_sort (c.axpy_like.begin (), c.axpy_like.end (),
                        [] (const std::pair<int, int64_t> &a, const std::pair<int, int64_t> &b2) { return a.second > b2.second; });
      double mean_task = c.axpy_like.empty () ? 0.0 : (double) c.total_values / (double) c.axpy_like.size ();
      c.bucket1 = bucket_of (mean_task);
      /* device allocations */
      size_t tb = c.tasks.size () * sizeof (bt_task);
      if (!rc) rc = hip_ok ("batch tasks", hipMalloc (&c.d_tasks, tb));
      if (!rc) rc = hip_ok ("batch tasks copy", hipMemcpy (c.d_tasks, c.tasks.data (), tb, hipMemcpyHostToDevice));
      if (!rc && ngroups > 0)
        {
          rc = hip_ok ("batch chains", hipMalloc (&c.d_chains, ch.size () * sizeof (bt_chain)));
          if (!rc) rc = hip_ok ("batch chains copy", hipMemcpy (c.d_chains, ch.data (), ch.size () * sizeof (bt_chain),
                                                                hipMemcpyHostToDevice));
          if (!rc) rc = hip_ok ("batch gbase", hipMalloc (&c.d_gbase, gbase.size () * sizeof (int64_t)));
          if (!rc) rc = hip_ok ("batch gbase copy", hipMemcpy (c.d_gbase, gbase.data (), gbase.size () * sizeof (int64_t),
                                                               hipMemcpyHostToDevice));
          const size_t lb = (size_t) scr_n * (nl + 1) * parts * 8;   /* records of nl limbs + sign/exponent */
          if (!rc) rc = hip_ok ("batch scratch", hipMalloc (&c.d_scr_l, lb ? lb : 8));
          b->scratch_bytes += lb;
        }
      if (!rc) rc = hip_ok ("batch counters", hipMalloc (&c.d_counters, 64));
      c.dev.tasks = (const bt_task *) c.d_tasks;
      c.dev.chains = (const bt_chain *) c.d_chains;
      c.dev.gbase = (const i64 *) c.d_gbase;
      c.dev.ngroups = ngroups;
      c.dev.scr = (u64 *) c.d_scr_l;
      c.dev.scr_n = scr_n;
      c.dev.counters = (unsigned *) c.d_counters;
      c.dev.pb = c.pb;
      c.dev.cplx = c.cplx;
      if (!rc) rc = upload_items (c, c.map1.chunk, s);
      if (rc) break;
    }
  /* ---- phase-2 units of all classes, longest first ---- */
  b->map2 = DEFAULT_P2;
  if (!rc)
    {
      std::vector<std::pair<int64_t, bt_unit>> lu;
      int64_t maxlen = 0;
      for (int ci = 0; ci < (int) b->classes.size (); ci++)
        {
          const auto &c = b->classes[ci];
          for (int64_t g = 0; g < (int64_t) c.gmax.size (); g++)
            for (int part = 0; part <= c.cplx; part++)
              {
                bt_unit u;
                u.cls = ci; u.part = part; u.fg = g;
                lu.push_back ({c.gmax[g], u});
                maxlen = std::max (maxlen, c.gmax[g]);
              }
        }
      std::stable_sort (lu.begin (), lu.end (),
                        [] (const std::pair<int64_t, bt_unit> &x, const std::pair<int64_t, bt_unit> &y) { return x.first > y.first; });
      for (auto &e : lu) b->units.push_back (e.second);
      b->bucket_units = bucket_of ((double) b->units.size ());
      b->bucket_len = bucket_of ((double) maxlen);
      rc = hip_ok ("batch counter", hipMalloc (&b->d_counter, 64));
      if (!rc && !b->units.empty ())
        {
          rc = hip_ok ("batch units", hipMalloc (&b->d_units, b->units.size () * sizeof (bt_unit)));
          if (!rc) rc = hip_ok ("batch units copy", hipMemcpy (b->d_units, b->units.data (), b->units.size () * sizeof (bt_unit),
                                                               hipMemcpyHostToDevice));
        }
      if (!rc) rc = hip_ok ("batch classes", hipMalloc (&b->d_classes, std::max<size_t> (1, b->classes.size ()) * sizeof (bt_class)));
    }
  /* the class table for the fold kernel (after tuning: items may be re-made) */
  auto upload_classes = [&] () -> int {
    std::vector<bt_class> tab;
    for (auto &c : b->classes) tab.push_back (c.dev);
    if (tab.empty ()) return (int) MPGFX_OK;
    return hip_ok ("batch classes copy", hipMemcpy (b->d_classes, tab.data (), tab.size () * sizeof (bt_class),
                                                    hipMemcpyHostToDevice));
  };
  if (!rc) rc = upload_classes ();
  /* ---- tuning ---- */
  if (!rc && tuning)
    {
      std::map<std::string, mpgfx_mapping> cmap = read_cache (cache);
      bool dirty = false;
      void *tune_out = nullptr;
      rc = redirect_axpy_outputs (b, true, &tune_out);
      for (auto &c : b->classes)
        if (!rc) rc = tune_class (c, b->cus, s, cmap, &dirty);
      if (!rc) rc = upload_classes ();
      if (!rc) rc = tune_folds (b, s, cmap, &dirty);
      if (!rc) rc = hip_ok ("batch tuning sync", hipStreamSynchronize (s));
      const int rrc = redirect_axpy_outputs (b, false, &tune_out);
      rc = rc ? rc : rrc;
      if (!rc && dirty) write_cache (cache, cmap);
    }
  /* ---- graph: per class [reset -> products] on its own branch, then the
   * folds of all classes in one kernel; 
   * captured on a private stream (the caller's may be the legacy null stream,
   * which cannot be captured); the graph is launched on the caller's stream ---- */
  if (!rc) rc = hip_ok ("batch stream sync", hipStreamSynchronize (s));
  if (!rc)
    {
      hipStream_t cs = nullptr;
      hipEvent_t fork = nullptr;
      std::vector<hipEvent_t> jv (b->classes.size (), nullptr);
      rc = hip_ok ("batch capture stream", hipStreamCreateWithFlags (&cs, hipStreamNonBlocking));
      if (!rc) rc = hip_ok ("batch event", hipEventCreateWithFlags (&fork, hipEventDisableTiming));
      for (size_t i = 0; !rc && i < b->classes.size (); i++)
        {
          hipStream_t ss;
          rc = hip_ok ("batch stream", hipStreamCreateWithFlags (&ss, hipStreamNonBlocking));
          if (!rc) b->side.push_back (ss);
          if (!rc) rc = hip_ok ("batch event", hipEventCreateWithFlags (&jv[i], hipEventDisableTiming));
        }
      if (!rc) rc = hip_ok ("batch capture", hipStreamBeginCapture (cs, hipStreamCaptureModeThreadLocal));
      if (!rc)
        {
          int lrc = hip_ok ("batch fork", hipEventRecord (fork, cs));
          for (size_t i = 0; !lrc && i < b->classes.size (); i++)
            {
              lrc = hip_ok ("batch fork wait", hipStreamWaitEvent (b->side[i], fork, 0));
              if (!lrc) lrc = launch_class (b->classes[i], b->cus, b->side[i]);
              if (!lrc) lrc = hip_ok ("batch join", hipEventRecord (jv[i], b->side[i]));
              if (!lrc) lrc = hip_ok ("batch join wait", hipStreamWaitEvent (cs, jv[i], 0));
            }
          if (!lrc) lrc = launch_folds (b, b->map2, cs);
          hipGraph_t g = nullptr;
          int erc = hip_ok ("batch end capture", hipStreamEndCapture (cs, &g));
          rc = lrc ? lrc : erc;
          if (!rc) b->graph = g;
          else if (g) (void) hipGraphDestroy (g);
        }
      if (!rc) rc = hip_ok ("batch instantiate", hipGraphInstantiate (&b->exec, b->graph, nullptr, nullptr, 0));
      if (fork) (void) hipEventDestroy (fork);
      for (auto &e : jv)
        if (e) (void) hipEventDestroy (e);
      if (cs) (void) hipStreamDestroy (cs);
    }
  if (rc)
    {
      mpgfx_batch_destroy (b);
      return rc;
    }
  *out = b;
  return MPGFX_OK;
}

extern "C" int mpgfx_batch_run (mpgfx_batch *b, void *stream)
{
  if (!b || !b->exec) return MPGFX_EINVAL;
  hipError_t e = hipGraphLaunch (b->exec, (hipStream_t) stream);
  if (e != hipSuccess) { set_error ("mpgfx_batch_run", e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

extern "C" int mpgfx_batch_info (const mpgfx_batch *b, size_t *scratch_bytes, int *classes)
{
  if (!b) return MPGFX_EINVAL;
  if (scratch_bytes) *scratch_bytes = b->scratch_bytes;
  if (classes) *classes = (int) b->classes.size ();
  return MPGFX_OK;
}
