// Batched 1D FFT along one axis of a distributed array, with the layout
// change of a reshape fused into the same pass.
//
// The reference's slab/pencil drivers pair FFTW guru stages over local axes
// with runReshapePlan between them (fftw_test.c:298-336,
// fftw_bench.c:261-305).  dm_plan_axis_fft does one such stage and the
// reshape that follows it in one device pass: read the input decomposition,
// transform along `axis`, write the output decomposition.
#include <algorithm>
#include <cstring>

#include "internal.h"
#include "fft_program.h"

struct dm_axis_fft {
  Decomp in = nullptr, out = nullptr;  // aliased
  int axis = 0, dbl = 1, inverse = 0;
  // compiled for these buffers
  const void *in_ptr = nullptr;
  void *out_ptr = nullptr;
  size_t in_slot = 0, out_slot = 0;
  std::vector<dm::FftLaunch> passes;
  void *mem = nullptr;
};

extern "C" struct dm_axis_fft *dm_plan_axis_fft(Decomp in, Decomp out, int axis, int precision, int direction) {
  if (in == nullptr || out == nullptr) {
    dm::fail(nullptr, "dm_plan_axis_fft: NULL decomposition");
    return nullptr;
  }
  if (compatible(in, out) != COMPATIBLE || axis < 0 || axis >= in->ndim) {
    raise_error(in->comm, "dm_plan_axis_fft: decompositions not compatible or axis out of range");
    return nullptr;
  }
  if (direction != -1 && direction != 1) {
    raise_error(in->comm, "dm_plan_axis_fft: direction must be FFTW_FORWARD or FFTW_BACKWARD");
    return nullptr;
  }
  std::string why;
  if (!dm::layout_rank_map_ok(in, &why) || !dm::layout_rank_map_ok(out, &why)) {
    raise_error(in->comm, "dm_plan_axis_fft: %s", why.c_str());
    return nullptr;
  }
  dm_axis_fft *p = new dm_axis_fft;
  p->in = in;
  p->out = out;
  p->axis = axis;
  p->dbl = precision ? 1 : 0;
  p->inverse = direction == 1 ? 1 : 0;
  return p;
}

extern "C" void dm_free_axis_fft(struct dm_axis_fft *p) {
  if (p == nullptr) return;
  if (p->mem) dm::device_free(p->mem);
  delete p;
}

extern "C" int dm_run_axis_fft(struct dm_axis_fft *p, const void *in, void *out) {
  if (p == nullptr) return DM_ERR_ARG;
  dm_comm_t comm = p->in->comm;
  dm::check_device(comm);
  size_t es = p->dbl ? 16 : 8;
  dm::Buffer bi, bo;
  size_t oi, oo;
  if (!dm::resolve(in, &bi, &oi) || !dm::resolve(out, &bo, &oo)) {
    raise_error(comm, "dm_run_axis_fft: buffers must be distributed device buffers (dm_malloc/dm_register)");
    return DM_ERR_ARG;
  }
  if (oi % es || oo % es || bi.slot % es || bo.slot % es) {
    raise_error(comm, "dm_run_axis_fft: buffers must be aligned to the element size");
    return DM_ERR_ARG;
  }
  if (p->mem == nullptr || p->in_ptr != in || p->out_ptr != out || p->in_slot != bi.slot ||
      p->out_slot != bo.slot) {
    if (p->mem) dm::device_free(p->mem);
    p->mem = nullptr;
    dm::HostLayout hi = dm::layout_from_decomp(p->in), ho = dm::layout_from_decomp(p->out);
    if (bi.nranks < hi.proc_used || bo.nranks < ho.proc_used) {
      raise_error(comm, "dm_run_axis_fft: buffers have fewer rank slots than the decompositions use");
      return DM_ERR_ARG;
    }
    size_t ri = (bi.slot - oi) / es, ro = (bo.slot - oo) / es;
    if (!dm::layout_usable(comm, "input", hi, ri) || !dm::layout_usable(comm, "output", ho, ro)) return DM_ERR_ARG;
    if (bi.base == bo.base) {
      size_t ei = 0, eo = 0;
      for (auto e : hi.slot_extent) ei = std::max(ei, (size_t)e * es);
      for (auto e : ho.slot_extent) eo = std::max(eo, (size_t)e * es);
      bool overlap = oi < oo + eo && oo < oi + ei;
      // in place is safe only when both sides are the same layout: every
      // tile then writes exactly the addresses it read
      if (overlap && !(oi == oo && equivalent(p->in, p->out))) {
        raise_error(comm, "dm_run_axis_fft: input and output overlap with different layouts");
        return DM_ERR_ARG;
      }
    }
    dm::Side si = dm::make_side(hi, (i64)(bi.slot / es)), so = dm::make_side(ho, (i64)(bo.slot / es));
    dm::TableBuilder tb;
    std::vector<dm::PendingPass> pp(1);
    if (!dm::plan_pass(comm, si, so, p->axis, p->dbl, p->inverse, bi.base + oi, bo.base + oo, tb, &pp[0]))
      return DM_ERR_ARG;
    p->mem = dm::finalize_passes(comm, tb, pp, &p->passes);
    if (p->mem == nullptr) return DM_ERR_ARG;
    p->in_ptr = in;
    p->out_ptr = out;
    p->in_slot = bi.slot;
    p->out_slot = bo.slot;
  }
  void *s = dm::world_stream(comm->world);
  for (auto &L : p->passes)
    if (!dm::launch_ok(p->in->comm, dm::launch_fft_pass(L, s), "FFT pass launch")) return DM_ERR_ARG;
  dm::world_finish(comm->world);
  return DM_SUCCESS;
}
