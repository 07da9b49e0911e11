"""Filter parameters and synthetic-scene settings of the named workloads.

The filter keys are those the reference reads from its libconfig file
(cheesinglee/scphd-camera-calibration src/scphdcameracalibration.cu:827-1005).
Rig: 640x480 image (u0 = 320, v0 = 240, image = 2u0 x 2v0 as in
src/disparitymeasurementmodel.cu:159-160), focal length f with the
reference's sign convention fx = fy = -f (config/calibration.cfg:74-75), so
that d = -fx/z = f/z > 0 in front of the camera.
"""
import math

DEG = math.pi / 180.0


def base_params(n_particles, samples, max_gaussians, lam, f=500.0, baseline=0.2, seed=1):
    """Filter parameters shared by every workload.

    Prior on the right-camera extrinsics: uniform over nominal +- 3 sigma with
    sigma = (0.02 m, 1 deg); rates start at zero.  Birth terms sit at the
    measurement with d = f / 4.5 m (the middle of the 2-10 m depth range)
    and wide disparity-rate variances.

    Two birth values decide whether the filter calibrates at all:

    * w0 = 1e-5 (the reference's calibration.cfg uses 1e-6).  The reference
      normalises with the birth/clutter term log(M w0 + kappa)
      (src/scphdcameracalibration.cu:1744-1745); with w0 = 1e-3 that term is
      ~log(0.2) at 200 measurements, above every detection term's bound
      log(pd / 2 pi), so no landmark weight ever grows past a few percent
      and the parent weights carry no information.
    * var_d0 = (f / 16)^2: the birth disparity is d0 +- f / 16, i.e. depth
      2.9-10 m at two sigma.  With (f / 8)^2, ~4 % of a birth cloud's samples
      have d < 0 (behind the camera); their non-detection clouds are out of
      view, keep weight ps^k, and live ~180 frames, which grows the maps
      past 1000 Gaussians per parent within 20 frames of the 30-landmark
      scene.

    The camera's constant-velocity model roughens the parents with
    accelerations of 0.002 m/s^2 and 0.1 deg/s^2: the rates it adds are not
    observable and random-walk, so twice these values let the estimate of a
    calibrated filter drift off again (half of eight scene seeds), while
    these keep it (all eight).

    With these values the oracle's rotation error falls over the frames of
    the 30-landmark scene and the best parent's map weight approaches the
    number of visible landmarks (tests/test_calibration.py).
    """
    sx, sa = 0.02, 1.0 * DEG
    nominal = [baseline, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    half = [3 * sx] * 3 + [0.0] * 3 + [3 * sa] * 3 + [0.0] * 3
    return {
        "n_particles": n_particles,
        "particles_per_feature": samples,
        "max_gaussians": max_gaussians,
        "storage": 0,
        "verbosity": 0,
        "slam": 0,
        "max_steps": 100000,
        "save_all_maps": 0,
        "min_weight": 1e-4,
        "min_separation": 4.0,
        "min_neff": 0.75,
        "ps": 0.95,
        "prior_a": [n - h for n, h in zip(nominal, half)],
        "prior_b": [n + h for n, h in zip(nominal, half)],
        "cam_accel_std": [0.002, 0.002, 0.002, 0.1 * DEG, 0.1 * DEG, 0.1 * DEG],
        "feat_accel_std": [0.02, 0.02, 0.02],
        "fx": -f, "fy": -f, "u0": 320.0, "v0": 240.0,
        "std_u": 1.0, "std_v": 1.0, "pd": 0.95, "lambda": float(lam),
        "birth_w0": 1e-5,
        "birth_d0": f / 4.5, "birth_vu0": 0.0, "birth_vv0": 0.0, "birth_vd0": 0.0,
        "birth_var_d0": (f / 16.0) ** 2, "birth_var_vu0": 25.0, "birth_var_vv0": 9.0, "birth_var_vd0": 1.0,
        "max_velocity": 5.0, "min_range": 0.5,
        "seed": seed,
    }


# name -> (filter parameters, scene settings)
WORKLOADS = {
    # Plain one-rig, small: 30 landmarks, lambda 10, 256 parents, S = 100,
    # cap 500 Gaussians per parent, 50 frames.
    "c0": dict(
        params=base_params(256, 100, 500, 10.0),
        scene=dict(n_landmarks=30, lam=10.0, f=500.0, baseline=0.2, n_frames=50, dt=0.5),
    ),
    # Plain one-rig, large: 200 landmarks, lambda 50, 4096 parents, S = 200,
    # cap 2000, 100 frames.
    "c1": dict(
        params=base_params(4096, 200, 2000, 50.0),
        scene=dict(n_landmarks=200, lam=50.0, f=500.0, baseline=0.2, n_frames=100, dt=0.5),
    ),
    # Precision-study size: 16384 parents, S = 100, 300 landmarks, lambda 100.
    "c3": dict(
        params=base_params(16384, 100, 2000, 100.0),
        scene=dict(n_landmarks=300, lam=100.0, f=500.0, baseline=0.2, n_frames=200, dt=0.5),
    ),
}


def workload(name, **overrides):
    w = WORKLOADS[name]
    p = dict(w["params"])
    s = dict(w["scene"])
    for k, v in overrides.items():
        if k in p:
            p[k] = v
        elif k in s:
            s[k] = v
        else:
            raise KeyError(k)
    return p, s


def rig(seed, n_particles, n_landmarks, lam, f=500.0, baseline=0.2, samples=100, max_gaussians=0,
        n_frames=100, dt=0.5, scene_rig=0):
    """One rig of a batched or served workload: filter parameters keyed by
    `seed` and the settings of its synthetic scene (scene.generate with
    seed=seed, rig=scene_rig)."""
    p = base_params(n_particles, samples, max_gaussians, lam, f=f, baseline=baseline, seed=seed)
    s = dict(n_landmarks=n_landmarks, lam=float(lam), f=float(f), baseline=float(baseline), n_frames=n_frames,
             dt=dt, rig=scene_rig)
    return p, s


def _splitmix64(master, index):
    x = (master + 0x9E3779B97F4A7C15 * (index + 1)) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def c2_rigs(master=1, n_rigs=64, samples=100, n_frames=100):
    """Batched ragged workload: n_rigs independent rigs, each with parents
    uniform in {256, 512, 1024, 2048, 4096}, landmarks uniform in [20, 300],
    clutter lambda uniform in [5, 80], baseline uniform in [0.1, 0.5] m and
    f uniform in [400, 800] px; rig r's filter seed and scene draw come from
    SplitMix64(master, r).  S (unspecified by the workload) = `samples`."""
    out = []
    for r in range(n_rigs):
        sd = _splitmix64(master, r)
        u = [((_splitmix64(sd, j) >> 11) * 2.0 ** -53) for j in range(5)]
        n_p = [256, 512, 1024, 2048, 4096][min(int(u[0] * 5), 4)]
        out.append(rig(seed=sd, n_particles=n_p, n_landmarks=20 + int(u[1] * 281), lam=5.0 + 75.0 * u[2],
                       baseline=0.1 + 0.4 * u[3], f=400.0 + 400.0 * u[4], samples=samples, n_frames=n_frames,
                       scene_rig=r))
    return out


def small_ragged_rigs(master=1, n_rigs=64, samples=100, n_frames=40):
    """The C2 draw at small sizes, for the ensemble leg: parents uniform in
    {256, 512}, landmarks uniform in [20, 100], clutter lambda uniform in
    [5, 40], baseline uniform in [0.1, 0.5] m, f uniform in [400, 800] px;
    seeds and scene draws from SplitMix64(master, r)."""
    out = []
    for r in range(n_rigs):
        sd = _splitmix64(master, r)
        u = [((_splitmix64(sd, j) >> 11) * 2.0 ** -53) for j in range(5)]
        out.append(rig(seed=sd, n_particles=[256, 512][min(int(u[0] * 2), 1)], n_landmarks=20 + int(u[1] * 81),
                       lam=5.0 + 35.0 * u[2], baseline=0.1 + 0.4 * u[3], f=400.0 + 400.0 * u[4], samples=samples,
                       n_frames=n_frames, scene_rig=r))
    return out


def c4_arrivals(master=1, rate=20.0, horizon=30.0, samples=100, fps=30.0, dur=(2.0, 10.0), parents=(256, 8192),
                landmarks=(20, 400), lam=(5.0, 100.0)):
    """Served workload: rigs arriving as a Poisson process of `rate` rigs/s
    over `horizon` seconds; parents log-uniform in `parents`, landmarks
    uniform in `landmarks`, clutter lambda uniform in `lam`, frames at `fps`
    for a duration uniform in `dur`.  Returns [(arrival_s, params, scene)]."""
    out, t, r = [], 0.0, 0
    while True:
        sd = _splitmix64(master, r)
        u = [((_splitmix64(sd, j) >> 11) * 2.0 ** -53) for j in range(6)]
        t += -math.log(1.0 - u[0]) / rate
        if t > horizon:
            return out
        n_p = int(round(math.exp(math.log(parents[0]) + u[1] * (math.log(parents[1]) - math.log(parents[0])))))
        n_frames = max(1, int((dur[0] + (dur[1] - dur[0]) * u[2]) * fps))
        p, s = rig(seed=sd, n_particles=n_p, n_landmarks=landmarks[0] + int(u[3] * (landmarks[1] - landmarks[0] + 1)),
                   lam=lam[0] + (lam[1] - lam[0]) * u[4], samples=samples, n_frames=n_frames, dt=1.0 / fps,
                   scene_rig=r)
        out.append((t, p, s))
        r += 1


def to_cfg(p):
    """Render parameters as a libconfig file in the schema the reference code
    reads (prior.{cartesian,angular}.{x..vz}.{a,b})."""
    comp = ["x", "y", "z", "vx", "vy", "vz"]
    lines = [
        f"verbosity = {p['verbosity']}",
        f"save_all_maps = {'true' if p['save_all_maps'] else 'false'}",
        f"slam = {'true' if p['slam'] else 'false'}",
        f"max_steps = {p['max_steps']}",
        f"n_particles = {p['n_particles']}",
        f"particles_per_feature = {p['particles_per_feature']}",
        f"max_gaussians = {p['max_gaussians']}",
        f"min_weight = {p['min_weight']!r}",
        f"min_separation = {p['min_separation']!r}",
        f"max_velocity = {p['max_velocity']!r} ;",
        f"min_range = {p['min_range']!r} ;",
        f"min_neff = {p['min_neff']!r} ;",
        f"ps = {p['ps']!r}",
        f"seed = {p['seed']}",
        "prior : {",
    ]
    for grp, base in (("cartesian", 0), ("angular", 6)):
        lines.append(f"    {grp} : {{")
        for i, c in enumerate(comp):
            lines.append(f"        {c} : {{a = {p['prior_a'][base + i]!r} ; b = {p['prior_b'][base + i]!r}}};")
        lines.append("    }")
    lines.append("}")
    ca = p["cam_accel_std"]
    lines += [
        "camera_motion_model : {",
        f"    cartesian : {{ ax = {ca[0]!r}\n        ay = {ca[1]!r}\n        az = {ca[2]!r} }}",
        f"    angular : {{ ax = {ca[3]!r}\n        ay = {ca[4]!r}\n        az = {ca[5]!r} }}",
        "}",
        "feature_motion_model : {",
        f"    ax = {p['feat_accel_std'][0]!r}",
        f"    ay = {p['feat_accel_std'][1]!r}",
        f"    az = {p['feat_accel_std'][2]!r}",
        "}",
        "measurement_model : {",
    ]
    for k in ("fx", "fy", "u0", "v0", "std_u", "std_v"):
        lines.append(f"    {k} = {p[k]!r}")
    lines.append(f"    lambda = {p['lambda']!r}")
    lines.append(f"    pd = {p['pd']!r}")
    lines.append("}")
    lines.append("birth : {")
    for k in ("w0", "d0", "vu0", "vv0", "vd0", "var_d0", "var_vu0", "var_vv0", "var_vd0"):
        lines.append(f"    {k} = {p['birth_' + k]!r} ;")
    lines.append("}")
    return "\n".join(lines) + "\n"
