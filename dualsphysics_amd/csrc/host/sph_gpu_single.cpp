// Host driver of the SPH step: a re-implementation of the reference's
// single-GPU solver flow (JSphGpuSingle / JSphGpu / JCellDivGpuSingle) on top
// of the C ABI in dsph_abi.h. It owns the particle arrays (same SoA layout),
// derives the constants like JSph::ConfigConstants1/2 and the cell map like
// JSph::ConfigCellDivision / SelecDomain / ConfigPosCellGpu, and runs
// ComputeStep_Sym / ComputeStep_Ver + RunCellDivide in the reference order.
//
// This file talks to the device only through dsph_* entry points.
#include "sph_gpu_single.h"

// SymplecticDtPre=min(dt_p,dt_c), time+=dt and the dt used, on the device.
int dsph_sim::dsph_dt_variable_min_advance() { return dsph_dt_advance(ctx, d_dt); }

extern "C" {

int dsph_sim_create(int device, const dsph_case* cs, uint32_t np, const uint32_t* h_idp,
                    const double* h_pos, const float* h_velrho, const float* h_boundnor,
                    dsph_sim** out) {
  if (!out) return DSPH_E_ARG;
  *out = nullptr;
  if (!cs || !h_idp || !h_pos || !h_velrho || !np || np >= DSPH_NP_LIMIT) return DSPH_E_ARG;
  dsph_sim* s = new dsph_sim();
  int rc = dsph_ctx_create(device, &s->ctx);
  if (rc) {
    delete s;
    return rc;
  }
  try {
    s->config(*cs);
    s->load(np, h_idp, h_pos, h_velrho, h_boundnor);
  } catch (const Err& e) {
    s->err = e.what();
    *out = s;
    return e.code;
  }
  *out = s;
  return DSPH_OK;
}

int dsph_sim_set_time(dsph_sim* sim, double timestep, double symplecticdtpre, uint32_t nstep) {
  if (!sim || !(timestep >= 0)) return DSPH_E_ARG;
  sim->TimeStep = timestep;
  if (symplecticdtpre > 0) sim->SymplecticDtPre = symplecticdtpre;
  sim->Nstep = nstep;
  return DSPH_OK;
}

int dsph_sim_destroy(dsph_sim* sim) {
  delete sim;
  return DSPH_OK;
}

const char* dsph_sim_last_error(dsph_sim* sim) { return sim ? sim->err.c_str() : "null sim"; }

int dsph_sim_run(dsph_sim* sim, uint32_t nsteps, int sync_dt) {
  if (!sim) return DSPH_E_ARG;
  try {
    sim->run(nsteps, sync_dt != 0);
  } catch (const Err& e) {
    sim->err = e.what();
    return e.code;
  }
  return DSPH_OK;
}

int dsph_sim_stats(dsph_sim* sim, dsph_stats* o) {
  if (!sim || !o) return DSPH_E_ARG;
  o->np = sim->Np; o->npb = sim->Npb; o->npbok = sim->NpbOk; o->nct = sim->Nct;
  o->nstep = sim->Nstep; o->npf_out_total = sim->NpfOutTotal;
  o->timestep = sim->TimeStep; o->dt_last = sim->DtLast;
  o->velmax = sim->VelMax; o->acemax = sim->AceMax; o->viscdtmax = sim->ViscDtMax;
  o->ndiv = sim->Ndiv; o->ndivfull = sim->NdivFull;
  o->symplecticdtpre = sim->SymplecticDtPre;
  return DSPH_OK;
}

int dsph_sim_download(dsph_sim* sim, uint32_t* np, uint32_t* h_idp, double* h_pos,
                      float* h_velrho, uint16_t* h_code) {
  if (!sim) return DSPH_E_ARG;
  try {
    const size_t n = sim->Np;
    if (np) *np = sim->Np;
    if (h_idp) sim->chk(dsph_memcpy_d2h(sim->ctx, h_idp, sim->idp, 4 * n), "download idp");
    if (h_velrho) sim->chk(dsph_memcpy_d2h(sim->ctx, h_velrho, sim->velrho, 16 * n), "download velrho");
    if (h_code) sim->chk(dsph_memcpy_d2h(sim->ctx, h_code, sim->code, 2 * n), "download code");
    if (h_pos) {
      std::vector<double> xy(2 * n), z(n);
      sim->chk(dsph_memcpy_d2h(sim->ctx, xy.data(), sim->posxy, 16 * n), "download posxy");
      sim->chk(dsph_memcpy_d2h(sim->ctx, z.data(), sim->posz, 8 * n), "download posz");
      for (size_t p = 0; p < n; p++) {
        h_pos[3 * p] = xy[2 * p]; h_pos[3 * p + 1] = xy[2 * p + 1]; h_pos[3 * p + 2] = z[p];
      }
    }
  } catch (const Err& e) {
    sim->err = e.what();
    return e.code;
  }
  return DSPH_OK;
}

int dsph_sim_download_forces(dsph_sim* sim, float* h_ace, float* h_ar) {
  if (!sim) return DSPH_E_ARG;
  try {
    const size_t n = sim->Np;
    if (h_ace) sim->chk(dsph_memcpy_d2h(sim->ctx, h_ace, sim->ace, 12 * n), "download ace");
    if (h_ar) sim->chk(dsph_memcpy_d2h(sim->ctx, h_ar, sim->ar, 4 * n), "download ar");
  } catch (const Err& e) {
    sim->err = e.what();
    return e.code;
  }
  return DSPH_OK;
}

int dsph_sim_set_closure(dsph_sim* sim, const dsph_mlp* m) {
  if (!sim) return DSPH_E_ARG;
  try {
    if (!m) {
      sim->UseClosure = false;
      return DSPH_OK;
    }
    if (sim->TVisco != DSPH_VISCO_LAMINARSPS)
      throw Err(DSPH_E_STATE, "the learned closure needs Laminar+SPS viscosity");
    sim->Closure = *m;
    if (!sim->velgrad) {
      sim->velgrad = sim->alloc<float>(9 * size_t(sim->NpCap));
      sim->closure_nu = sim->alloc<float>(size_t(sim->NpCap));
      sim->chk(dsph_memset_async(sim->ctx, sim->closure_nu, 0, 4 * size_t(sim->NpCap)), "nu");
    }
    sim->UseClosure = true;
  } catch (const Err& e) {
    sim->err = e.what();
    return e.code;
  }
  return DSPH_OK;
}

int dsph_sim_closure_nu(dsph_sim* sim, float* h_nu, float* h_velgrad) {
  if (!sim || !h_nu) return DSPH_E_ARG;
  if (!sim->closure_nu) {
    sim->err = "dsph_sim_closure_nu: the learned closure was never enabled";
    return DSPH_E_STATE;
  }
  int rc = dsph_memcpy_d2h(sim->ctx, h_nu, sim->closure_nu, 4 * size_t(sim->Np));
  if (!rc && h_velgrad)
    rc = dsph_memcpy_d2h(sim->ctx, h_velgrad, sim->velgrad, 36 * size_t(sim->Np));
  if (rc) sim->err = dsph_last_error(sim->ctx);
  return rc;
}

int dsph_sim_filter_probes(dsph_sim* sim, uint32_t nprobe, const double* h_probes, double hf,
                           double* h_out) {
  if (!sim || !h_probes || !h_out) return DSPH_E_ARG;
  if (!nprobe) return DSPH_OK;
  double *dp = nullptr, *dout = nullptr;
  int rc = DSPH_OK;
  try {
    dp = sim->alloc<double>(3 * size_t(nprobe));
    dout = sim->alloc<double>(23 * size_t(nprobe));
    sim->chk(dsph_memcpy_h2d(sim->ctx, dp, h_probes, 24 * size_t(nprobe)), "probes");
    const dsph_divdata d = sim->divdata();
    sim->chk(dsph_filter_probes(sim->ctx, nprobe, dp, hf, &d, sim->MassFluid, sim->posxy,
                                sim->posz, sim->velrho, sim->code, dout),
             "filter probes");
    sim->chk(dsph_memcpy_d2h(sim->ctx, h_out, dout, 8 * 23 * size_t(nprobe)), "probe moments");
  } catch (const Err& e) {
    sim->err = e.what();
    rc = e.code;
  }
  dsph_free(sim->ctx, dp);
  dsph_free(sim->ctx, dout);
  return rc;
}

int dsph_sim_gauge_points(dsph_sim* sim, uint32_t n, const double* h_points, float* h_vel,
                          float* h_mass) {
  if (!sim || !h_points || (!h_vel && !h_mass)) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  double* dp = nullptr;
  float *dv = nullptr, *dm = nullptr;
  int rc = DSPH_OK;
  try {
    dp = sim->alloc<double>(3 * size_t(n));
    if (h_vel) dv = sim->alloc<float>(3 * size_t(n));
    if (h_mass) dm = sim->alloc<float>(n);
    sim->chk(dsph_memcpy_h2d(sim->ctx, dp, h_points, 24 * size_t(n)), "gauge points");
    const dsph_divdata d = sim->divdata();
    sim->chk(dsph_gauge_points(sim->ctx, n, dp, &d, sim->MapRealPosMin, sim->MapRealPosMax,
                               sim->posxy, sim->posz, sim->velrho, sim->code, dv, dm),
             "gauge points");
    if (h_vel) sim->chk(dsph_memcpy_d2h(sim->ctx, h_vel, dv, 12 * size_t(n)), "gauge vel");
    if (h_mass) sim->chk(dsph_memcpy_d2h(sim->ctx, h_mass, dm, 4 * size_t(n)), "gauge mass");
  } catch (const Err& e) {
    sim->err = e.what();
    rc = e.code;
  }
  dsph_free(sim->ctx, dp);
  dsph_free(sim->ctx, dv);
  dsph_free(sim->ctx, dm);
  return rc;
}

int dsph_sim_gauge_maxz(dsph_sim* sim, uint32_t n, const double* h_points,
                        const float* h_height_dist, float* h_zmax) {
  if (!sim || !h_points || !h_height_dist || !h_zmax) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  double* dp = nullptr;
  float *dh = nullptr, *dz = nullptr;
  int rc = DSPH_OK;
  try {
    dp = sim->alloc<double>(3 * size_t(n));
    dh = sim->alloc<float>(2 * size_t(n));
    dz = sim->alloc<float>(n);
    sim->chk(dsph_memcpy_h2d(sim->ctx, dp, h_points, 24 * size_t(n)), "maxz points");
    sim->chk(dsph_memcpy_h2d(sim->ctx, dh, h_height_dist, 8 * size_t(n)), "maxz sizes");
    const dsph_divdata d = sim->divdata();
    sim->chk(dsph_gauge_maxz(sim->ctx, n, dp, dh, &d, sim->MapRealPosMin[2], sim->posxy,
                             sim->posz, sim->code, dz),
             "gauge maxz");
    sim->chk(dsph_memcpy_d2h(sim->ctx, h_zmax, dz, 4 * size_t(n)), "gauge zmax");
  } catch (const Err& e) {
    sim->err = e.what();
    rc = e.code;
  }
  dsph_free(sim->ctx, dp);
  dsph_free(sim->ctx, dh);
  dsph_free(sim->ctx, dz);
  return rc;
}

int dsph_sim_gauge_force(dsph_sim* sim, uint32_t idbegin, uint32_t count, float* h_ace) {
  if (!sim || !h_ace) return DSPH_E_ARG;
  if (!count) return DSPH_OK;
  float* da = nullptr;
  int rc = DSPH_OK;
  try {
    da = sim->alloc<float>(3 * size_t(count));
    const dsph_divdata d = sim->divdata();
    sim->chk(dsph_gauge_force(sim->ctx, sim->Npb, idbegin, count, &d, sim->posxy, sim->posz,
                              sim->code, sim->idp, sim->velrho, da),
             "gauge force");
    sim->chk(dsph_memcpy_d2h(sim->ctx, h_ace, da, 12 * size_t(count)), "gauge force ace");
  } catch (const Err& e) {
    sim->err = e.what();
    rc = e.code;
  }
  dsph_free(sim->ctx, da);
  return rc;
}

int dsph_sim_pips(dsph_sim* sim, uint64_t h_res[5]) {
  if (!sim || !h_res) return DSPH_E_ARG;
  const dsph_divdata d = sim->divdata();
  const int rc = dsph_pips_count_split_sync(sim->ctx, sim->NpbOk, 0, sim->Np - sim->Npb, sim->Npb, &d,
                                      sim->dcell, sim->poscell, h_res);
  if (rc) sim->err = dsph_last_error(sim->ctx);
  return rc;
}

int dsph_sim_floating(dsph_sim* sim, uint32_t cf, dsph_floating_state* out) {
  if (!sim || !out) return DSPH_E_ARG;
  if (cf >= sim->ft.size()) {
    sim->err = "dsph_sim_floating: no floating body " + std::to_string(cf);
    return DSPH_E_ARG;
  }
  const auto& f = sim->ft[cf];
  for (int k = 0; k < 3; k++) {
    out->center[k] = f.center[k];
    out->angles[k] = f.angles[k]; out->fvel[k] = f.fvel[k]; out->fomega[k] = f.fomega[k];
    out->facelin[k] = f.facelin[k]; out->faceang[k] = f.faceang[k];
  }
  return DSPH_OK;
}

int dsph_sim_interact(dsph_sim* sim) {
  if (!sim) return DSPH_E_ARG;
  try {
    sim->pre_interaction();
    sim->interaction();
    sim->chk(dsph_synchronize(sim->ctx), "synchronize");
  } catch (const Err& e) {
    sim->err = e.what();
    return e.code;
  }
  return DSPH_OK;
}

dsph_ctx* dsph_sim_ctx(dsph_sim* sim) { return sim ? sim->ctx : nullptr; }

}  // extern "C"
