"""Learned initial guess for RHF on a family of geometries of one molecule.

A 3-layer perceptron (width 256, BF16 weights, FP32 accumulation; libqcml.so kernels for both
training and inference) maps per-atom-pair distance descriptors (1/r_ij over all atom pairs, in the
molecule's fixed atom order) to the correction that takes the SAD density to the converged one, in
the Loewdin-orthogonalised AO basis: dP = S^{1/2} (D - D_SAD) S^{1/2} (upper triangle).  The SAD
part carries the atoms' own densities, which move with the atoms; the network learns the bonding
rearrangement.  At inference P = S^{1/2} D_SAD S^{1/2} + dP is symmetrised and purified: the nocc
eigenvectors of P with the largest eigenvalues give D = 2 S^{-1/2} U U^T S^{-1/2}, an idempotent
density with the right electron count, which seeds the SCF in place of SAD or the core guess.  Training targets are converged FP64 densities that the engine itself produces (batched RHF
from the SAD seed).  No reference counterpart: the reference seeds with the core guess or SAD
(ChargeDensity/Imp/Seed.C:64-77).
"""
import os
import time

import numpy as np

import qchem_amd as qa
from qchem_amd import systems

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "guess_data", "c3h6o_mlp.npz")


def descriptor(R):
    R = np.asarray(R, dtype=np.float64).reshape(-1, 3)
    i, j = np.triu_indices(len(R), 1)
    return 1.0 / np.linalg.norm(R[i] - R[j], axis=1)


def lowdin(S):
    s, U = np.linalg.eigh(S)
    return (U * np.sqrt(s)) @ U.T, (U / np.sqrt(s)) @ U.T


def pack_upper(P):
    return P[np.triu_indices(P.shape[0])]


def unpack_upper(v, n):
    P = np.zeros((n, n))
    P[np.triu_indices(n)] = v
    return P + np.triu(P, 1).T


def purify(P, S_mhalf, nocc):
    """Idempotent density with nocc doubly occupied orbitals closest to the predicted P."""
    w, U = np.linalg.eigh(0.5 * (P + P.T))
    Uo = U[:, -nocc:]
    C = S_mhalf @ Uo
    return 2.0 * C @ C.T


def overlap(mol):
    B = qa.Basis(mol["shells"])
    S, _, _ = qa.JKEngine(B).one_electron(mol["Z"], mol["R"])
    return B, S


def make_dataset(indices, chunk=128, log=print, **scf):
    """Converged FP64 densities (batched RHF, SAD seed) for acetone_like(i), i in indices."""
    X, Y, E, ok = [], [], [], []
    scf = dict(dict(e_tol=1e-7, fd_tol=1e-4, max_iter=80, jk_flags=qa.QCJK_J | qa.QCJK_K | qa.QCJK_SCREEN,
                    screen_tol=1e-11), **scf)
    for c0 in range(0, len(indices), chunk):
        t0 = time.time()
        mols = [systems.acetone_like(i) for i in indices[c0:c0 + chunk]]
        bases = [qa.Basis(m["shells"]) for m in mols]
        out, summ = qa.rhf_batch(bases, mols, guess=qa.GUESS_SAD, **scf)
        eng = qa.JKEngine(bases)
        for k, (m, r, b) in enumerate(zip(mols, out, bases)):
            S, _, _ = eng.one_electron(m["Z"], m["R"], mol=k)
            Sh, _ = lowdin(S)
            X.append(descriptor(m["R"]))
            Y.append(pack_upper(Sh @ (r["D"] - qa.sad_density(b, m["Z"], m["R"])) @ Sh))
            E.append(r["energy"])
            ok.append(r["converged"])
        log("dataset chunk %d-%d: %d/%d converged, %d batched builds, %.1f s" % (
            indices[c0], indices[min(c0 + chunk, len(indices)) - 1], summ["nconverged"], len(mols),
            summ["iterations"], time.time() - t0))
    return np.array(X), np.array(Y), np.array(E), np.array(ok)


class LearnedGuess:
    """Trained network + normalisation; ``density(mol)`` returns the purified seed density."""

    def __init__(self, net, x_mean, x_std, y_mean, y_scale, nbf, nocc):
        self.net, self.x_mean, self.x_std = net, x_mean, x_std
        self.y_mean, self.y_scale, self.nbf, self.nocc = y_mean, y_scale, nbf, nocc

    @classmethod
    def train(cls, X, Y, nbf, nocc, width=256, epochs=400, batch=128, lr=1e-3, seed=7, nval=64, log=print):
        """Adam with cosine decay on the MSE of the normalised orthogonalised density; the last nval
        samples are held back as a validation set, checked every few epochs, and the parameters with
        the lowest validation loss are kept (early stopping)."""
        import torch
        from qchem_amd import mlp
        Xv, Yv = X[-nval:], Y[-nval:]
        X, Y = X[:-nval], Y[:-nval]
        x_mean, x_std = X.mean(0), X.std(0) + 1e-12
        y_mean = Y.mean(0)
        y_scale = float((Y - y_mean).std()) + 1e-12
        Xn = torch.from_numpy(((X - x_mean) / x_std).astype(np.float32)).cuda()
        Yn = torch.from_numpy(((Y - y_mean) / y_scale).astype(np.float32)).cuda()
        Xvn = torch.from_numpy(((Xv - x_mean) / x_std).astype(np.float32)).cuda()
        Yvn = ((Yv - y_mean) / y_scale).astype(np.float32)
        net = mlp.Net(X.shape[1], width, Y.shape[1], max_batch=max(batch, nval), seed=seed)
        out = torch.empty((nval, Y.shape[1]), dtype=torch.float32, device="cuda")

        def val_loss():
            net.forward(Xvn, out)
            return float(np.mean((out.cpu().numpy() - Yvn) ** 2))
        n = X.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        steps = epochs * max(1, n // batch)
        t0, k = time.time(), 0
        best, best_p, check = float("inf"), None, max(1, epochs // 40)
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).cuda()
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                rate = lr * 0.5 * (1.0 + np.cos(np.pi * k / steps))  # cosine decay
                want = (b0 == 0 and (ep % max(1, epochs // 8) == 0 or ep == epochs - 1))
                loss = net.train_step(Xn[idx], Yn[idx], rate, want_loss=want)
                if want:
                    log("epoch %d batch loss %.5f validation loss %.5f (%.1f s)" % (ep, loss, val_loss(),
                                                                                   time.time() - t0))
                k += 1
            if ep % check == check - 1 or ep == epochs - 1:
                v = val_loss()
                if v < best:
                    best, best_p = v, net.get_params()
        net.set_params(best_p)
        log("kept the parameters with validation loss %.5f" % best)
        return cls(net, x_mean, x_std, y_mean, y_scale, nbf, nocc)

    def predict_orth(self, mols):
        import torch
        Xn = np.array([(descriptor(m["R"]) - self.x_mean) / self.x_std for m in mols], dtype=np.float32)
        out = torch.empty((len(mols), self.y_mean.size), dtype=torch.float32, device="cuda")
        self.net.forward(torch.from_numpy(Xn).cuda(), out)
        return out.cpu().numpy().astype(np.float64) * self.y_scale + self.y_mean

    def density(self, mol):
        B, S = overlap(mol)
        Sh, Smh = lowdin(S)
        P = Sh @ qa.sad_density(B, mol["Z"], mol["R"]) @ Sh + unpack_upper(self.predict_orth([mol])[0], self.nbf)
        return purify(P, Smh, self.nocc)

    def save(self, path=DEFAULT_MODEL):
        p = self.net.get_params()
        nin, width, nout = self.net.dims
        os.makedirs(os.path.dirname(path), exist_ok=True)
        # BF16 bits of every parameter (the kernels multiply with BF16 weights) + the FP32 biases
        u = p.view(np.uint32).astype(np.uint64)
        bits = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
        biases = np.concatenate([p[o:o + k] for o, k in _bias_slices(nin, width, nout)])
        np.savez_compressed(path, dims=np.array([nin, width, nout]), w_bf16=bits, b_f32=biases,
                            x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean.astype(np.float32),
                            y_scale=np.array([self.y_scale]), nbf=np.array([self.nbf]), nocc=np.array([self.nocc]))

    @classmethod
    def load(cls, path=DEFAULT_MODEL, max_batch=512):
        from qchem_amd import mlp
        z = np.load(path, allow_pickle=False)
        nin, width, nout = (int(x) for x in z["dims"])
        net = mlp.Net(nin, width, nout, max_batch=max_batch)
        p = (z["w_bf16"].astype(np.uint32) << 16).view(np.float32).copy()
        acc = 0
        for o, k in _bias_slices(nin, width, nout):
            p[o:o + k] = z["b_f32"][acc:acc + k]
            acc += k
        net.set_params(p)
        return cls(net, z["x_mean"], z["x_std"], z["y_mean"].astype(np.float64), float(z["y_scale"][0]),
                   int(z["nbf"][0]), int(z["nocc"][0]))


def _bias_slices(nin, width, nout):
    """(offset, length) of b1, b2, b3 in the flat parameter vector W1 b1 W2 b2 W3 b3."""
    return [(nin * width, width), (nin * width + width + width * width, width),
            (nin * width + 2 * width + width * width + width * nout, nout)]


def time_to(e_hist, t_hist, e_ref, tol=1e-6):
    for e, t in zip(e_hist, t_hist):
        if abs(e - e_ref) < tol:
            return float(t)
    return None


def iters_to(e_hist, e_ref, tol=1e-6):
    for k, e in enumerate(e_hist):
        if abs(e - e_ref) < tol:
            return k + 1
    return None


FAST = dict(incremental=8, screen_tol=1e-12)  # incremental builds of the density change, Schwarz-screened


def evaluate(model, indices, log=print, mixed_tol=1e-7):
    """Per held-out geometry: time to 1e-6 Eh of the converged FP64 energy for
    sad                       plain FP64 full builds from the SAD seed (the baseline)
    learned_fp64 / _mixed     the learned seed with FP64 / mixed-precision full builds
    sad_incremental           FP64 incremental screened builds (FAST) from SAD
    learned_mixed_incremental the learned seed, mixed precision and FAST builds (the optimised engine)
    The learned timings include the seed's inference."""
    rows = []
    tight = dict(e_tol=1e-10, fd_tol=1e-7, max_iter=150)
    warm = systems.acetone_like(indices[0])
    qa.rhf(qa.Basis(warm["shells"]), warm["Z"], warm["R"], guess=qa.GUESS_SAD, max_iter=2)  # fills the SAD cache
    for i in indices:
        m = systems.acetone_like(i)
        B = qa.Basis(m["shells"])
        sad = qa.rhf(B, m["Z"], m["R"], guess=qa.GUESS_SAD, **tight)
        e_conv = sad["energy"]
        t0 = time.perf_counter()
        D0 = model.density(m)
        t_inf = time.perf_counter() - t0
        row = {"index": i, "e_conv": e_conv, "sad_converged": sad["converged"], "inference_s": t_inf,
               "sad": {"iterations": sad["iterations"], "time_to_1e-6_s": time_to(sad["energy_history"],
                                                                                    sad["time_history"], e_conv),
                       "iterations_to_1e-6": iters_to(sad["energy_history"], e_conv),
                       "first_energy_error": float(sad["energy_history"][0] - e_conv)}}
        mixed = dict(jk_flags=qa.QCJK_J | qa.QCJK_K | qa.QCJK_MIXED, mixed_tol=mixed_tol)
        fast = dict(FAST, jk_flags=qa.QCJK_J | qa.QCJK_K | qa.QCJK_SCREEN)
        fast_mixed = dict(FAST, jk_flags=qa.QCJK_J | qa.QCJK_K | qa.QCJK_SCREEN | qa.QCJK_MIXED, mixed_tol=mixed_tol)
        for name, kw, seeded in (("learned_fp64", {}, True), ("learned_mixed", mixed, True),
                                 ("sad_incremental", fast, False), ("learned_mixed_incremental", fast_mixed, True)):
            if seeded:
                r = qa.rhf(B, m["Z"], m["R"], D0=D0, **tight, **kw)
            else:
                r = qa.rhf(B, m["Z"], m["R"], guess=qa.GUESS_SAD, **tight, **kw)
            tt = time_to(r["energy_history"], r["time_history"], e_conv)
            if tt is not None and seeded:
                tt += t_inf
            row[name] = {"iterations": r["iterations"], "converged": r["converged"], "energy_diff": r["energy"] - e_conv,
                         "time_to_1e-6_s": tt,
                         "iterations_to_1e-6": iters_to(r["energy_history"], e_conv),
                         "first_energy_error": float(r["energy_history"][0] - e_conv)}
        rows.append(row)
        log("held-out %d: to 1e-6 Eh SAD %s it %.3f s (first error %.2e)  learned %s it %.3f s (first error %.2e)"
            "  learned+mixed+incremental %.3f s" % (
                i, row["sad"]["iterations_to_1e-6"], row["sad"]["time_to_1e-6_s"] or -1, row["sad"]["first_energy_error"],
                row["learned_fp64"]["iterations_to_1e-6"], row["learned_fp64"]["time_to_1e-6_s"] or -1,
                row["learned_fp64"]["first_energy_error"], row["learned_mixed_incremental"]["time_to_1e-6_s"] or -1))
    return rows


def summarize(rows):
    def med(key, sub):
        v = [r[key][sub] for r in rows if r[key][sub] is not None]
        return float(np.median(v)) if v else None
    s = {"n": len(rows)}
    for k in ("sad", "learned_fp64", "learned_mixed", "sad_incremental", "learned_mixed_incremental"):
        s[k] = {"median_iterations": med(k, "iterations"), "median_iterations_to_1e-6": med(k, "iterations_to_1e-6"),
                "median_time_to_1e-6_s": med(k, "time_to_1e-6_s"),
                "median_abs_first_energy_error": float(np.median([abs(r[k]["first_energy_error"]) for r in rows]))}
    base = s["sad"]["median_time_to_1e-6_s"]
    for k in ("learned_fp64", "learned_mixed", "sad_incremental", "learned_mixed_incremental"):
        if base and s[k]["median_time_to_1e-6_s"]:
            s["speedup_%s_vs_sad_fp64" % k] = base / s[k]["median_time_to_1e-6_s"]
        s[k]["max_abs_energy_diff"] = max(abs(r[k]["energy_diff"]) for r in rows)
    return s


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train and evaluate the learned SCF guess on C3H6O-like geometries")
    ap.add_argument("--train", type=int, default=1024)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--report", default=None)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--save-data", default=None, help="also write the first 1024 samples (float16) here")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    import torch
    torch.zeros(1, device="cuda")  # torch's HIP runtime first: it serves as device-memory plumbing below
    t0 = time.time()
    tr = list(range(args.train))
    ho = list(range(100000, 100000 + args.heldout))  # geometries never seen in training
    X, Y, E, ok = make_dataset(tr, log=log)
    log("training set: %d geometries, %d converged, %.1f s" % (len(tr), int(ok.sum()), time.time() - t0))
    X, Y = X[ok], Y[ok]
    if args.save_data:
        np.savez_compressed(args.save_data, X=X.astype(np.float32), Y=Y[:1024].astype(np.float16))
    m0 = systems.acetone_like(0)
    nbf = systems.nbf(m0["shells"])
    model = LearnedGuess.train(X, Y, nbf, sum(m0["Z"]) // 2, epochs=args.epochs, batch=args.batch, log=log)
    model.save(args.out)
    log("saved %s" % args.out)
    rows = evaluate(model, ho, log=log)
    summ = summarize(rows)
    log(json.dumps(summ))
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"summary": summ, "rows": rows, "train": len(tr), "train_converged": int(ok.sum()),
                       "epochs": args.epochs}, f, indent=1)


if __name__ == "__main__":
    main()
