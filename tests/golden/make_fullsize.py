"""Runs the REFERENCE's own classes (oracle/_ref/ref_harness, compiled in
place from the reference tree) on the full-size drop-in configurations and
keeps only hashes, norms and subsampled fields (small) in tests/golden:

  D2Q9 256^2 periodic, tau 1, G -5.0, noisy rho 0.7 +- 0.01 (seed 1),
       10,000 steps (4 ranks)
  D3Q19 256^3 periodic, tau 1, G -5.5, droplet R 48, W 2, rho 2.0 / 0.15,
       2,000 steps (8 ranks)

    make -C oracle ref && python tests/golden/make_fullsize.py c1|c2
"""
import hashlib, json, os, subprocess, sys, tempfile, time
import numpy as np
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle
H = os.path.join(ROOT, 'oracle', '_ref', 'ref_harness')
def run(name, q, shape, tau, G, init, steps, ranks):
    nx, ny = shape[0], shape[1]; nz = shape[2] if len(shape) == 3 else 1
    n = nx * ny * nz
    rho0 = oracle.noise_rho(shape, init[1]) if init[0] == 'noise' else oracle.droplet_rho(shape, init[1])
    pre = os.path.join(tempfile.mkdtemp(), name)
    rho0.astype(np.float64).tofile(pre + '_rho0.bin')
    t0 = time.time()
    subprocess.run(['/opt/conda/bin/mpirun', '-np', str(ranks), H, str(q), str(nx), str(ny), str(nz), repr(tau), repr(G), '0', pre + '_rho0.bin', pre, str(steps)], check=True)
    dt = time.time() - t0
    a = np.fromfile(f'{pre}_{steps}.bin', dtype=np.float64)
    rho, u, v, w, f = a[:n], a[n:2*n], a[2*n:3*n], a[3*n:4*n], a[4*n:]
    sh = lambda x: hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()
    out = dict(name=name, q=q, shape=list(shape), tau=tau, G=G, init=list(init), steps=steps, ranks=ranks,
               seconds=dt, sha256_rho=sh(rho), sha256_u=sh(u), sha256_v=sh(v), sha256_w=sh(w), sha256_f=sh(f),
               norm_rho=float(np.linalg.norm(rho)), norm_u=float(np.linalg.norm(u)), sum_rho=float(rho.sum()),
               min_rho=float(rho.min()), max_rho=float(rho.max()),
               source='reference classes via oracle/_ref/ref_harness (mpirun -np %d)' % ranks)
    step = 8 if q == 19 else 4
    rs = rho.reshape(shape)
    sub = rs[::step, ::step, ::step] if q == 19 else rs[::step, ::step]
    np.savez_compressed(os.path.join(ROOT, 'tests', 'golden', f'full_{name}.npz'), rho_sub=sub, u_sub=(u.reshape(shape)[::step, ::step, ::step] if q == 19 else u.reshape(shape)[::step, ::step]),
                        meta=np.frombuffer(json.dumps(out).encode(), dtype=np.uint8))
    for fn in (pre + '_rho0.bin', f'{pre}_{steps}.bin'):
        os.remove(fn)
    print(json.dumps(out), flush=True)
if sys.argv[1] == 'c1':
    run('d2q9_256_spinodal_10000', 9, (256, 256), 1.0, -5.0, ('noise', 1), 10000, 4)
else:
    run('d3q19_256_droplet_2000', 19, (256, 256, 256), 1.0, -5.5, ('droplet', 48.0), 2000, 8)
