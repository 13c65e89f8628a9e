"""Runner: the training loop of exp_runner.py on supernormal_amd.

``Runner(conf_text, mode='train', is_continue=False, datadir=None)`` reads the
reference's HOCON keys (train.*, model.sdf_network, model.encoding,
model.variance_network, model.ray_marching.*, val.report_freq/save_freq) and
``train()`` runs the iteration of exp_runner.py:147-210:
  step size 10^(log10(start) - slope*i)          (exp_runner.py:81-83, 149)
  occupancy update every n steps                 (152-155)
  bandwidth +1 every increase_bindwidth_every    (158-159)
  random 3x3 patches, near/far on the unit sphere (162-167)
  render + normal/eikonal/mask losses            (177-203)
  backward + Adam, warm-up + cosine learning rate (205-210, 270-279)

``fused=True`` (default) runs the iteration on supernormal_amd.fused.FusedStep;
``fused=False`` runs the reference's structure (NeuSRenderer.render + torch
autograd over the drop-in kernels).  Both share one Adam kernel over a flat
parameter buffer.

Data: the reference loads DiLiGenT-MV from disk (dataset_loader.py); this
runner takes a synthetic scene named by ``dataset.synthetic`` (sphere128,
tbs612, prims1024; see supernormal_amd.scene) or a Dataset object.

Differences from the reference, on purpose: an iteration with no surviving
sample is skipped (the reference raises KeyError, exp_runner.py:182); the
checkpoint also stores the occupancy grid, bandwidth and RNG state (the
reference's resume loses them, exp_runner.py:306-315); occupancy resolution
and max samples per ray come from the config (renderer.py:50 hard-codes 128).
"""
from __future__ import annotations

import math
import os
import time
from typing import Optional

import numpy as np
import torch
import torch.nn.functional as F

from .conf import Conf, parse
from .fields import SDFNetwork, SingleVarianceNetwork
from .fused import FlatParams, FusedStep
from .renderer import NeuSRenderer


class Runner:
    def __init__(self, conf_text, mode="train", is_continue=False, datadir=None, fused=True,
                 dataset=None, device="cuda", base_exp_dir=None, seed=0, rng_seed=None):
        self.device = torch.device(device)
        self.conf_text = conf_text if isinstance(conf_text, str) else None
        self.conf = conf_text if isinstance(conf_text, dict) else parse(conf_text)
        if not isinstance(self.conf, Conf):
            self.conf = Conf(self.conf)
        c = self.conf
        torch.manual_seed(seed)
        self.base_exp_dir = base_exp_dir or c.get("general.base_exp_dir", "./exp/run")
        if dataset is None:
            dataset = self._config_dataset(c, seed, datadir)
        self.dataset = dataset
        self.iter_step = 0
        self.end_iter = c.get_int("train.end_iter")
        self.batch_size = c.get_int("train.batch_size")
        self.patch_size = c.get_int("train.patch_size", 3)
        self.learning_rate = c.get_float("train.learning_rate")
        self.learning_rate_alpha = c.get_float("train.learning_rate_alpha")
        self.warm_up_end = c.get_float("train.warm_up_end", 0.0)
        self.loss_type = c.get("train.loss_type", "l1")
        self.normal_weight = c.get_float("train.normal_weight")
        self.eikonal_weight = c.get_float("train.eikonal_weight")
        self.mask_weight = c.get_float("train.mask_weight")
        self.increase_bindwidth_every = c.get_int("train.increase_bindwidth_every", 350)
        self.gradient_method = c.get("train.gradient_method", "dfd")
        self.report_freq = c.get_int("val.report_freq", 100)
        self.save_freq = c.get_int("val.save_freq", 10 ** 9)
        self.start_step_size = c.get_float("model.ray_marching.start_step_size", 1e-2)
        self.end_step_size = c.get_float("model.ray_marching.end_step_size", 5e-4)
        self.slop_step = (np.log10(self.start_step_size) - np.log10(self.end_step_size)) / self.end_iter
        self.occ_threshold = c.get_float("model.ray_marching.occ_threshold", 0.1)
        self.occ_update_freq = c.get_int("model.ray_marching.occ_update_freq", 8)
        self.occ_resolution = c.get_int("model.ray_marching.occ_resolution", 128)
        self.max_samples_per_ray = c.get_int("model.ray_marching.max_samples_per_ray", 0)
        # extensions: MLP precision of the fused kernels ('fp32' | 'fp16') and the
        # iteration driver ('fused': FusedStep; 'sync_free' / 'graph': SyncFreeStep,
        # eager or replayed as one HIP graph)
        self.mlp_precision = c.get("train.mlp_precision", "fp32")
        self.step_mode = c.get("train.step_mode", "fused")
        if self.step_mode not in ("fused", "sync_free", "graph"):
            raise ValueError(f"unknown train.step_mode {self.step_mode!r}")
        self.sdf_network = SDFNetwork(**c["model.sdf_network"],
                                      encoding_config=c["model.encoding"]).to(self.device)
        self.sdf_network.mlp_precision = self.mlp_precision
        self.deviation_network = SingleVarianceNetwork(**c["model.variance_network"]).to(self.device)
        self.renderer = NeuSRenderer(self.sdf_network, self.deviation_network,
                                     self.gradient_method, occ_resolution=self.occ_resolution,
                                     device=self.device)
        params = list(self.sdf_network.parameters()) + list(self.deviation_network.parameters())
        self.optimizer = FlatParams(params, self.device)
        self.lr = self.learning_rate
        self.fused = fused
        self.fused_step = FusedStep(self.sdf_network, self.deviation_network,
                                    self.renderer.occupancy_grid, self.gradient_method,
                                    self.max_samples_per_ray, self.normal_weight,
                                    self.eikonal_weight, self.mask_weight, self.loss_type)
        self.rng_seed = rng_seed
        self.sf_step = None
        if fused and self.step_mode != "fused":
            self._make_sync_free_step()
        if rng_seed is not None:
            # patch sampling and occupancy sampling from generators of this run only
            g = torch.Generator(device=self.device)
            g.manual_seed(int(rng_seed) + 1)
            self.renderer.occupancy_grid.generator = g
        self.mode = mode
        self.history = []
        if is_continue:
            self.load_checkpoint()

    def _config_dataset(self, c, seed, datadir=None):
        """The dataset the config names (exp_runner.py:45-48): `dataset.synthetic`
        (one of supernormal_amd.scene's workloads) or the reference's on-disk
        format (`dataset.data_dir` etc., supernormal_amd.dataset_loader), with
        `general.dataset_class` naming the reference's loader when given."""
        ds_conf = c.get("dataset")
        if not isinstance(ds_conf, dict):
            raise KeyError("the config has no dataset block")
        cls = c.get("general.dataset_class")
        if "synthetic" in ds_conf:
            from .scene import config_dataset
            return config_dataset(ds_conf["synthetic"], self.device, seed)
        if cls not in (None, "models.dataset_loader.Dataset",
                       "supernormal_amd.dataset_loader.Dataset"):
            raise ValueError(f"general.dataset_class {cls!r} is not supported (the reference's "
                             "models.dataset_loader.Dataset format or dataset.synthetic)")
        from .dataset_loader import load_dataset
        ds_conf = Conf(ds_conf)
        if datadir is not None:
            ds_conf["data_dir"] = datadir
        return load_dataset(ds_conf, self.device)

    def _make_sync_free_step(self):
        from .step import SyncFreeStep
        self.sf_step = SyncFreeStep(
            self.sdf_network, self.deviation_network, self.renderer.occupancy_grid,
            self.optimizer, self.dataset, self.batch_size, self.patch_size,
            self.max_samples_per_ray, self.normal_weight, self.eikonal_weight,
            self.mask_weight, self.loss_type, precision=self.mlp_precision,
            device=self.device, seed=self.rng_seed, method=self.gradient_method)

    # ------------------------------------------------------------------ train
    def update_learning_rate(self):
        if self.iter_step < self.warm_up_end:
            factor = self.iter_step / self.warm_up_end
        else:
            alpha = self.learning_rate_alpha
            progress = (self.iter_step - self.warm_up_end) / (self.end_iter - self.warm_up_end)
            factor = (np.cos(np.pi * progress) + 1.0) * 0.5 * (1 - alpha) + alpha
        self.lr = self.learning_rate * factor

    def occ_eval_fn(self):
        return self.fused_step.occ_eval_fn if self.fused else self.renderer.occ_eval_fn

    def begin_iteration(self, iter_i):
        """The host part of an iteration before its step (exp_runner.py:147-159):
        step size, occupancy update when due, bandwidth; returns the step size."""
        r = self.renderer
        r.sampling_step_size = 10 ** (np.log10(self.start_step_size) - self.slop_step * iter_i)
        r.occupancy_grid.every_n_step(step=iter_i, occ_eval_fn=self.occ_eval_fn(),
                                      occ_thre=self.occ_threshold, n=self.occ_update_freq)
        if self.iter_step % self.increase_bindwidth_every == 0:
            self.sdf_network.increase_bandwidth()
        return r.sampling_step_size

    def end_iteration(self):
        """After a device iteration that another driver replayed."""
        self.iter_step += 1
        self.update_learning_rate()

    def train_iteration(self, iter_i, ahead=None):
        """One iteration; returns the loss tensor (or None when skipped).
        In graph mode the next iteration's march may run ahead (see
        _march_ahead_step); `ahead` False forbids it, True allows it past
        end_iter (a caller that will run more iterations)."""
        r = self.renderer
        self.begin_iteration(iter_i)
        if self.fused and self.sf_step is not None:
            graph = self.step_mode == "graph"
            losses = self.sf_step.step(self.lr, r.sampling_step_size, use_graph=graph,
                                       next_step_size=self._march_ahead_step(iter_i, ahead)
                                       if graph else None)
            self.iter_step += 1
            self.update_learning_rate()
            return losses[:3].sum()
        if self.sf_step is not None:
            # leaving the device iteration: its step count is the optimiser's
            self.sf_step.drain()
            self.sf_step.sync_step_count()
        o, d, plane_n, v_inv, n_true, mask = self.dataset.gen_random_patches(
            self.batch_size, patch_H=self.patch_size, patch_W=self.patch_size)
        pc = self.patch_size // 2
        near, far = self.dataset.near_far_from_sphere(o[:, pc, pc], d[:, pc, pc])
        self.optimizer.zero_grad()
        if self.fused:
            res = self.fused_step.forward_backward(o, d, plane_n, v_inv, n_true, mask, near, far,
                                                   float(r.sampling_step_size))
            loss = res.loss
            self.last_samples = res.n_samples
        else:
            loss = self._reference_structured_loss(o, d, plane_n, v_inv, n_true, mask, near, far)
            if loss is not None:
                loss.backward()
        if loss is None:
            self.update_learning_rate()
            return None
        self.optimizer.collect()
        self.optimizer.adam(self.lr)
        self.iter_step += 1
        self.update_learning_rate()
        return loss

    def _march_ahead_step(self, iter_i, ahead=None):
        """The next iteration's step size when its march may run ahead of this
        iteration's end: not before an occupancy update, a bandwidth increase,
        a checkpoint or the end of training (each changes what the march reads
        or should see the state between iterations)."""
        if ahead is False:
            return None
        k = self.iter_step + 1  # iter_step once this iteration is done
        if ((iter_i + 1) % self.occ_update_freq == 0 or k % self.increase_bindwidth_every == 0
                or k % self.save_freq == 0 or (ahead is None and k >= self.end_iter)):
            return None
        return 10 ** (np.log10(self.start_step_size) - self.slop_step * (iter_i + 1))

    def _reference_structured_loss(self, o, d, plane_n, v_inv, n_true, mask, near, far):
        out = self.renderer.render(o, d, plane_n, near, far, v_inv,
                                   max_samples_per_ray=self.max_samples_per_ray)
        if out["gradients"] is None:
            return None
        self.last_samples = int(out["gradients"].shape[0])
        if self.mask_weight > 0.0:
            mask = (mask > 0.5).float()
        else:
            mask = torch.ones_like(mask)
        mask_sum = mask.sum() + 1e-5
        err = (out["comp_normal"] - n_true) * mask
        if self.loss_type == "l1":
            normal_loss = F.l1_loss(err, torch.zeros_like(err), reduction="sum") / mask_sum
        else:
            normal_loss = F.mse_loss(err, torch.zeros_like(err), reduction="sum") / mask_sum
        gn = torch.linalg.norm(out["gradients"], ord=2, dim=-1)
        eik = F.mse_loss(gn, torch.ones_like(gn), reduction="mean")
        mloss = F.binary_cross_entropy(out["weight_sum"].clip(1e-5, 1.0 - 1e-5), mask)
        return self.normal_weight * normal_loss + self.mask_weight * mloss + \
            self.eikonal_weight * eik

    def train(self, n_iters: Optional[int] = None, log=print):
        """The loop of exp_runner.py:145-220 for the remaining iterations (or
        n_iters).  The schedules (step size, occupancy updates) follow the
        absolute iteration, so a resumed run continues them; the reference
        restarts them at 0 on resume (its loop index, exp_runner.py:146-155)."""
        self.update_learning_rate()
        res_step = (self.end_iter - self.iter_step) if n_iters is None else n_iters
        first = self.iter_step
        t0 = time.time()
        for k in range(res_step):
            iter_i = first + k
            loss = self.train_iteration(iter_i)
            if loss is not None and self.iter_step % self.report_freq == 0:
                lv = float(loss)
                self.history.append((self.iter_step, lv))
                if log:
                    log(f"iter {self.iter_step}: loss {lv:.4e} s "
                        f"{float(self.deviation_network.variance):.3e} step "
                        f"{self.renderer.sampling_step_size:.3e} "
                        f"({time.time() - t0:.1f}s)")
            if self.iter_step % self.save_freq == 0 and self.iter_step > 0:
                self.save_checkpoint()

    # -------------------------------------------------------------- checkpoint
    def _ckpt_path(self, it=None):
        it = self.iter_step if it is None else it
        return os.path.join(self.base_exp_dir, "checkpoints", f"ckpt_{it:06d}.pth")

    def save_checkpoint(self):
        if self.sf_step is not None:
            self.sf_step.sync_step_count()
        os.makedirs(os.path.join(self.base_exp_dir, "checkpoints"), exist_ok=True)
        ck = {
            "sdf_network_fine": self.sdf_network.state_dict(),
            "variance_network_fine": self.deviation_network.state_dict(),
            "optimizer": {"exp_avg": self.optimizer.exp_avg.cpu(),
                          "exp_avg_sq": self.optimizer.exp_avg_sq.cpu(),
                          "step": self.optimizer.step_count},
            "iter_step": self.iter_step,
            "bindwidth": self.sdf_network.bindwidth,
            "occupancy_grid": self.renderer.occupancy_grid.state_dict(),
            "rng_cpu": torch.get_rng_state(),
            "rng_cuda": torch.cuda.get_rng_state(self.device) if self.device.type == "cuda" else None,
            "rng_sampler": self.sf_step.sampler_rng_state() if self.sf_step is not None else None,
            "rng_occupancy": (self.renderer.occupancy_grid.generator.get_state()
                              if self.renderer.occupancy_grid.generator is not None else None),
        }
        if self.sf_step is not None and self.sf_step.gen is None:
            # the sampler draws from torch's default generator
            ck["rng_cuda"] = ck["rng_sampler"]
        torch.save(ck, self._ckpt_path())

    def load_checkpoint(self, path=None):
        if path is None:
            d = os.path.join(self.base_exp_dir, "checkpoints")
            names = sorted(n for n in os.listdir(d) if n.endswith(".pth")
                           and int(n[5:-4]) <= self.end_iter)
            path = os.path.join(d, names[-1])
        # on the host: the RNG states must stay CPU byte tensors; the rest is
        # copied into the device buffers below
        ck = torch.load(path, map_location="cpu", weights_only=True)
        self.sdf_network.load_state_dict(ck["sdf_network_fine"])
        self.deviation_network.load_state_dict(ck["variance_network_fine"])
        # parameters were re-pointed at the flat buffer; copy values back into it
        params = list(self.sdf_network.parameters()) + list(self.deviation_network.parameters())
        self.optimizer = FlatParams(params, self.device)
        self.optimizer.exp_avg.copy_(ck["optimizer"]["exp_avg"].to(self.device))
        self.optimizer.exp_avg_sq.copy_(ck["optimizer"]["exp_avg_sq"].to(self.device))
        self.optimizer.step_count = int(ck["optimizer"]["step"])
        self.iter_step = int(ck["iter_step"])
        self.sdf_network.bindwidth = int(ck["bindwidth"])
        self.renderer.occupancy_grid.load_state_dict(ck["occupancy_grid"])
        torch.set_rng_state(ck["rng_cpu"])
        if ck.get("rng_cuda") is not None:
            torch.cuda.set_rng_state(ck["rng_cuda"], self.device)
        if ck.get("rng_occupancy") is not None and self.renderer.occupancy_grid.generator is not None:
            self.renderer.occupancy_grid.generator.set_state(ck["rng_occupancy"])
        if self.sf_step is not None:
            # new parameter buffers: new step buffers and graph
            self.sf_step.drain()
            torch.cuda.synchronize(self.device)
            self._make_sync_free_step()
            if ck.get("rng_sampler") is not None:
                self.sf_step.set_sampler_rng_state(ck["rng_sampler"])


def synthetic_conf(workload="tbs612", gradient_method="dfd", end_iter=None, lr=None,
                   batch_size=None, max_samples_per_ray=None, mlp_precision="fp32",
                   step_mode="fused") -> Conf:
    """The synthetic training workloads as configs with the reference's keys.

    sphere128: sphere r 0.5, 8 views 128^2, hash 8 levels x 2, T 2^15, base 16,
               x1.5, MLP 32 x 1 hidden, 256 patches, occupancy 64^3, <= 32
               samples/ray, lr 1e-2, 200 iterations, step size 1e-2.
    tbs612:    torus + box + 3 spheres, 20 views 612x512, hash 16 x 2, T 2^19,
               base 16, x1.38, MLP 64 x 2 hidden, 2048 patches (18 432 rays),
               occupancy 128^3, bandwidth +1 / 250 its, sample capacity 1024
               per centre ray (unbounded in effect).
    prims1024: 8 primitives with ridges, 50 views 1024^2, hash T 2^21, 8192
               patches, occupancy 256^3, sample capacity 1024 per ray.
    """
    if workload == "sphere128":
        enc = dict(otype="HashGrid", n_levels=8, n_features_per_level=2, log2_hashmap_size=15,
                   base_resolution=16, per_level_scale=1.5)
        net = dict(d_out=1, d_in=3, d_hidden=32, n_layers=1, skip_in=[-1], bias=0.6,
                   geometric_init=True, weight_norm=True, input_concat=True)
        train = dict(learning_rate=1e-2, end_iter=200, batch_size=256, increase_bindwidth_every=25)
        rm = dict(occ_resolution=64, max_samples_per_ray=32)
    elif workload in ("tbs612", "prims1024"):
        big = workload == "prims1024"
        enc = dict(otype="HashGrid", n_levels=16, n_features_per_level=2,
                   log2_hashmap_size=21 if big else 19, base_resolution=16, per_level_scale=1.38)
        net = dict(d_out=1, d_in=3, d_hidden=64, n_layers=2, skip_in=[-1], bias=0.6,
                   geometric_init=True, weight_norm=True, input_concat=True)
        train = dict(learning_rate=5e-4, end_iter=10000 if big else 5000,
                     batch_size=8192 if big else 2048, increase_bindwidth_every=250)
        # capacity of samples per centre ray: the reference marches unbounded;
        # a 128 / 256 cap truncates rays before the surface once the step size
        # has decayed (40% of rays at 1e-3) and training diverges, so the cap is
        # set where it is (practically) never reached
        rm = dict(occ_resolution=256 if big else 128, max_samples_per_ray=1024)
    else:
        raise ValueError(workload)
    train.update(learning_rate_alpha=0.05, patch_size=3, warm_up_end=50, loss_type="l2",
                 normal_weight=1.0, eikonal_weight=1.0, mask_weight=1.0,
                 gradient_method=gradient_method, mlp_precision=mlp_precision,
                 step_mode=step_mode)
    if end_iter is not None and workload != "sphere128":
        train["end_iter"] = end_iter
    if lr is not None:
        train["learning_rate"] = lr
    if batch_size is not None:
        train["batch_size"] = batch_size
    if max_samples_per_ray is not None:
        rm["max_samples_per_ray"] = max_samples_per_ray
    # the reference's step-size decay 1e-2 -> 1e-3; the small workload keeps 1e-2:
    # its 32-sample cap spans only 0.032 of a ray at 1e-3, less than the occupied
    # shell in front of the surface, and training with the decay diverges (the
    # CPU oracle's Chamfer over 200 iterations: 0.053 -> 0.41 with the decay,
    # 0.014 -> 0.0098 with 1e-2)
    rm.update(start_step_size=1e-2, end_step_size=1e-2 if workload == "sphere128" else 1e-3,
              occ_threshold=0.1, occ_update_freq=8)
    return Conf(general=Conf(base_exp_dir=f"./exp/{workload}"),
                dataset=Conf(synthetic=workload),
                train=Conf(train), val=Conf(report_freq=100, save_freq=10 ** 9),
                model=Conf(sdf_network=Conf(net), variance_network=Conf(init_val=0.5),
                           ray_marching=Conf(rm), encoding=Conf(enc)))


def build_parser():
    """The command line of exp_runner.py:639-645 (same flags and defaults)."""
    import argparse
    p = argparse.ArgumentParser(description="SuperNormal training on supernormal_amd")
    p.add_argument("--conf", type=str, default="./confs/base.conf")
    p.add_argument("--mode", type=str, default="eval_normal")
    p.add_argument("--mcube_threshold", type=float, default=0.0)
    p.add_argument("--is_continue", default=False, action="store_true")
    p.add_argument("--gpu", type=int, default=0)
    p.add_argument("--obj_name", type=str, default="")
    return p


def read_conf_text(path, obj_name):
    """The config file's text with CASE_NAME replaced by the object's name
    (exp_runner.py:651-653)."""
    with open(path) as f:
        return f.read().replace("CASE_NAME", obj_name)


def main(argv=None):
    """exp_runner.py:633-657: parse the flags, select the GPU, build the
    Runner from the config file and train."""
    args = build_parser().parse_args(argv)
    torch.cuda.set_device(args.gpu)
    print(f"Running on the object: {args.obj_name}")
    conf_text = read_conf_text(args.conf, args.obj_name)
    runner = Runner(conf_text, args.mode, args.is_continue, device=f"cuda:{args.gpu}")
    runner.train()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
