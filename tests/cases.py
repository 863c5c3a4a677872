"""Flux configurations shared by the oracle and GPU parity tests.

Each case is a list of term dicts in the oracle's / the C ABI's field names;
the parameters are those the golden driver used (tests/golden/manifest.txt).
"""
import golden_io as gi

_, P = gi.manifest()
GAMMA, R = P["gamma"], P["R"]
VC = dict(kind=7, law=0, mu=P["mu_ref"], prandtl=P["prandtl"])
VP = dict(kind=7, law=1, mu=P["mu_ref"], T_ref=P["T_ref"], power=P["power"], prandtl=P["prandtl"])
PDC = dict(kind=3, sensor=0, sensor_param=0.7, eps_p=0.3, eps_T=0.2)
PDD = dict(kind=3, sensor=1, sensor_param=1e-3, eps_p=0.3, eps_T=0.2)
K2 = dict(kind=1, order=2)
# sgs_visc_t(law, wale_t(gas, cw, delta, prt)) (viscous_laws.h:174-216, subgrid_scale.h:24-86)
WALE = dict(sgs=1, sgs_cw=0.5, sgs_delta=0.3, sgs_prt=0.9)
K4 = dict(kind=1, order=4)

FLUX_CASES = {
    "keep2": [K2],
    "keep4": [K4],
    "totani": [dict(kind=2)],
    "visc_const": [VC],
    "visc_pow": [VP],
    "keep2_visc_pow": [K2, VP],
    "pdiss_const": [PDC],
    "pdiss_ducros": [PDD],
    "weno_rus": [dict(kind=4, split=0, smooth=1)],
    "weno_rus_lin": [dict(kind=4, split=0, smooth=0)],
    "weno_phys": [dict(kind=4, split=1, smooth=1)],
    "fweno": [dict(kind=5, smooth=1)],
    "fweno_lin": [dict(kind=5, smooth=0)],
    "first_order": [dict(kind=6)],
    # weno_fds_t<rusanov_fds_t>: face-reconstructed states, Rusanov flux difference
    "weno_fds": [dict(kind=9, smooth=1)],
    "weno_fds_lin": [dict(kind=9, smooth=0)],
    "keep2_pdiss_visc_const": [K2, PDD, VC],
    # visc_lr with a laminar law plus the WALE subgrid viscosity
    "visc_sgs_const": [dict(VC, **WALE)],
    "keep2_visc_sgs_pow": [K2, dict(VP, **WALE)],
    # hybrid_scheme_t: cent_keep<order> blended with WENO by a face sensor
    "hybrid_keep2_weno_ducros_full": [dict(kind=8, order=2, sensor=1, sensor_param=1e-3, split=0, smooth=1,
                                           diss_kind=4, full=1)],
    "hybrid_keep2_fweno_const_diss": [dict(kind=8, order=2, sensor=0, sensor_param=0.7, smooth=1,
                                           diss_kind=5, full=0)],
    "hybrid_keep4_weno_ducros_full_visc": [dict(kind=8, order=4, sensor=1, sensor_param=1e-3, split=0, smooth=1,
                                                diss_kind=4, full=1), VC],
}

# cases whose arithmetic involves pow(): equal to the reference within
# round-off, not bit for bit, on the GPU (device pow vs glibc pow)
USES_POW = {"visc_pow", "keep2_visc_pow", "keep2_visc_sgs_pow"}

# compositions the CPU oracle reproduces bit for bit against the reference but
# the GPU library refuses (SPD_ERR_UNSUPPORTED): its first hybrid_scheme_t
# kernel build was followed by a GPU memory fault (in an unrelated launch)
# whose cause was never established, so the library has no device code for
# it and these cases are skipped there
GPU_UNSUPPORTED = {k for k in FLUX_CASES if k.startswith("hybrid")}
GPU_UNSUPPORTED_REASON = ("hybrid_scheme_t is pinned on the CPU oracle only; the device library refuses it "
                          "(an earlier device build was followed by a memory fault of unestablished cause)")

G1 = dict(nb=(2, 3, 2), nc=(4, 8, 4), ng=2)
G2 = dict(nb=(2, 2, 2), nc=(8, 8, 8), ng=2)
