//! The headless `sim` command (flags as documented in docs/features.md:18-35
//! and NOTES-for-data.md:15-22; the reference snapshot ships no main.rs).
//! Goes into the reference's sim crate as `sim/src/bin/sim.rs`; with the
//! `hip` feature, `--cuda` (the default when a device is present) runs the
//! gfx950 backend through `Sim::new(.., BackendKind::Cuda)`.
//!
//!   sim --net file.ne2 [--script x.rhai] [--eval '...']... [--ms 1000]
//!       [--dt 0.1] [--seed 1] [--cuda | --cpu] [--int-weights]
//!   sim --bench N [--bench-k 100] [--wscale 1.0] [--w-inh 5.0]
//!       [--warm 100] [--ms 1000] [--noise 0.0] [--int-weights]
use anyhow::Context;
use clap::Parser;
use sim::{bench, script::Env, BackendKind, Sim, SimConfig};

#[derive(Parser, Debug)]
#[command(name = "sim", about = "neuro_edit v2 simulator (headless)")]
struct Args {
    /// Network file (.ne2 or .json) to load.
    #[arg(long)]
    net: Option<String>,
    /// Rhai script run against the loaded network.
    #[arg(long)]
    script: Option<String>,
    /// Rhai snippets run in order, in the script's scope.
    #[arg(long)]
    eval: Vec<String>,
    /// Simulated time to run, ms.
    #[arg(long, default_value_t = 1000.0)]
    ms: f32,
    /// Integration step, ms.
    #[arg(long)]
    dt: Option<f32>,
    #[arg(long, default_value_t = 1)]
    seed: u64,
    /// Use the GPU backend (the gfx950 library with the `hip` feature).
    #[arg(long, conflicts_with = "cpu")]
    cuda: bool,
    /// Use the CPU reference backend.
    #[arg(long)]
    cpu: bool,
    /// Fixed-point delay ring: bitwise reproducible runs.
    #[arg(long)]
    int_weights: bool,
    /// Benchmark: random network of N neurons (sim/src/bench.rs).
    #[arg(long)]
    bench: Option<usize>,
    #[arg(long, default_value_t = 100)]
    bench_k: usize,
    #[arg(long, default_value_t = 1.0)]
    wscale: f32,
    #[arg(long, default_value_t = 5.0)]
    w_inh: f32,
    /// Warm-up before the timed bench run, ms.
    #[arg(long, default_value_t = 100.0)]
    warm: f32,
    /// Poisson kick rate per neuron, Hz.
    #[arg(long, default_value_t = 0.0)]
    noise: f32,
}

fn backend(a: &Args) -> BackendKind {
    if a.cpu {
        BackendKind::Cpu
    } else if a.cuda || sim::cuda_available() {
        BackendKind::Cuda
    } else {
        BackendKind::Cpu
    }
}

fn main() -> anyhow::Result<()> {
    let a = Args::parse();
    let kind = backend(&a);
    if let Some(n) = a.bench {
        let dt = a.dt.unwrap_or(net_format::DEFAULT_DT_MS);
        let net = bench::bench_net(n, a.bench_k, a.seed, a.wscale, a.w_inh, dt);
        let cfg = SimConfig {
            dt_ms: dt,
            noise_rate_hz: a.noise,
            seed: a.seed,
            int_weights: a.int_weights,
            ..SimConfig::default()
        };
        let r = bench::run_bench(&net, kind, cfg, a.ms, a.warm);
        println!(
            "{n} neurons x {} synapses, {:?}: {:.3} s wall for {} ms ({:.2}x realtime), {:.1} Hz, {} spikes, {} MB device",
            net.synapses.len(),
            kind,
            r.wall_s,
            a.ms,
            r.realtime_x,
            r.rate_hz,
            r.spikes,
            r.gpu_mem_mb
        );
        return Ok(());
    }
    let path = a.net.as_deref().context("--net or --bench is required")?;
    let mut net = net_format::io::load(path).map_err(|e| anyhow::anyhow!("{path}: {e:?}"))?;
    if let Some(dt) = a.dt {
        net.meta.dt_ms = dt;
    }
    if a.script.is_some() || !a.eval.is_empty() {
        let mut env = Env::new(net, kind, a.seed)?;
        if a.int_weights {
            env.eval_str("sim.set_int_weights(1);").map_err(anyhow::Error::msg)?;
        }
        if let Some(s) = &a.script {
            env.eval_file(s).map_err(anyhow::Error::msg)?;
        }
        for e in &a.eval {
            env.eval_str(e).map_err(anyhow::Error::msg)?;
        }
        return Ok(());
    }
    let mut cfg = SimConfig::default();
    sim::meta_to_cfg(&net.meta, &mut cfg);
    cfg.seed = a.seed;
    cfg.int_weights = a.int_weights;
    let mut s = Sim::new(&net, cfg, kind)?;
    let t0 = std::time::Instant::now();
    s.run_ms(a.ms)?;
    s.synchronize();
    let wall = t0.elapsed().as_secs_f64();
    println!(
        "{} neurons, {:?}: {} ms in {:.3} s ({:.2}x realtime), {} spikes, {:.2} Hz",
        s.n(),
        kind,
        a.ms,
        wall,
        a.ms as f64 / 1000.0 / wall,
        s.spikes_total(),
        s.mean_rate_hz()
    );
    Ok(())
}
