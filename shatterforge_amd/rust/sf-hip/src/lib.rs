//! Binding of `libsf_hip` (the MI355X physics hot path, `include/sf.h`) for the reference engine's Rust host.
//!
//! `ffi` is the C ABI as an `extern "C"` block; the rest wraps it in owned handles that free their device memory on
//! drop and turn the library's negative return codes into `SfError` carrying `sf_last_error_message()`.
//! `legacy` re-exports the legacy engine's own entry points (src/lib.rs `init`, `step`, `explode`, `detonate`,
//! `n_structs`, `ptr_structs`, `ptr_stats`, tools/debug_exports.rs `dbg_*`) under their reference names, so the
//! wasm page's calls can be pointed at the GPU library one for one.
//!
//! (This crate is source only in this repository: there is no Rust toolchain in its build image.)

use core::ffi::{c_char, c_int, c_void};
use std::ffi::CStr;
use std::fmt;

pub mod ffi {
    use super::*;

    #[repr(C)]
    pub struct SfWorld {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfLat {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfLoad {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfServe {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfBricks {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfWarm {
        _p: [u8; 0],
    }
    #[repr(C)]
    pub struct SfNet {
        _p: [u8; 0],
    }

    /// `sf_answer`: one served query's answer
    #[repr(C)]
    #[derive(Clone, Copy, Debug, Default)]
    pub struct SfAnswer {
        pub id: u64,
        pub bricks: i32,
        pub standing: i32,
        pub pieces: i32,
        pub load_sweeps: i32,
        pub failed: [u32; 4],
        pub fragments: [f32; 25],
        pub load_resid: f64,
        pub latency_ms: f64,
        pub queue_ms: f64,
    }

    pub const SF_BODY_F: usize = 20;
    pub const SF_SHAPE_F: usize = 13;
    pub const SF_LAT_BODY_F: usize = 22;
    pub const SF_LAT_BRICK_F: usize = 13;

    extern "C" {
        pub fn sf_last_error() -> c_int;
        pub fn sf_last_error_message() -> *const c_char;
        pub fn sf_device_count() -> c_int;

        // box core (engine/src/phys.rs World, engine/src/wasm.rs exports)
        pub fn sf_world_create(body_cap: c_int, shape_cap: c_int, params: *const f32) -> *mut SfWorld;
        pub fn sf_world_destroy(w: *mut SfWorld);
        pub fn sf_world_load(w: *mut SfWorld, nbody: c_int, bodies: *const f32, nshape: c_int, shape_body: *const i32,
                             shapes: *const f32) -> c_int;
        pub fn sf_world_step(w: *mut SfWorld, dt: f32) -> c_int;
        pub fn sf_world_read_bodies(w: *mut SfWorld, out: *mut f32) -> c_int;
        pub fn sf_world_nbody(w: *mut SfWorld) -> c_int;
        pub fn sf_world_pose(w: *mut SfWorld) -> *const f32;
        pub fn sf_world_particles_capacity(w: *mut SfWorld, cap: c_int) -> c_int;
        pub fn sf_world_add_particles(w: *mut SfWorld, n: c_int, rec8: *const f32, slots: *mut i32) -> c_int;
        pub fn sf_world_kill_particle(w: *mut SfWorld, i: c_int) -> c_int;
        pub fn sf_world_blast_particles(w: *mut SfWorld, cx: f32, cy: f32, cz: f32, reach: f32, speed: f32) -> c_int;
        pub fn sf_world_wake_particles(w: *mut SfWorld) -> c_int;
        pub fn sf_world_read_particles(w: *mut SfWorld, out12: *mut f32) -> c_int;
        pub fn sf_world_particle_stats(w: *mut SfWorld, out4: *mut i32) -> c_int;
        pub fn sf_world_particle_ms(w: *mut SfWorld, ms: *mut f32) -> c_int;
        pub fn sf_pile_setup(n: u32, substeps: u32);
        pub fn sf_stack_setup(n: u32, substeps: u32, sleep: u32);
        pub fn sf_step(dt: f32) -> u32;
        pub fn sf_body_count() -> u32;
        pub fn sf_manifolds() -> u32;
        pub fn sf_points() -> u32;
        pub fn sf_max_speed() -> f32;
        pub fn sf_body_pos(i: u32, axis: u32) -> f32;
        pub fn sf_pose() -> *const f32;

        // legacy brick engine, handle API (src/lib.rs World)
        pub fn sf_lat_create(brick_cap: c_int, body_cap: c_int, max_scenes: c_int, colours: c_int, rounds: c_int) -> *mut SfLat;
        pub fn sf_lat_destroy(l: *mut SfLat);
        pub fn sf_lat_scene_add(l: *mut SfLat, brick_cap: c_int, body_cap: c_int, max_structs: c_int, seed: u32) -> c_int;
        pub fn sf_lat_add_structure(l: *mut SfLat, scene: c_int, n: c_int, xyz: *const i32, mat: *const u8,
                                    infill: *const u8) -> c_int;
        pub fn sf_lat_add_body(l: *mut SfLat, scene: c_int, n: c_int, xyz: *const i32, mat: *const u8, v3: *const f32,
                               w3: *const f32) -> c_int;
        pub fn sf_lat_finalize(l: *mut SfLat, scene: c_int) -> c_int;
        pub fn sf_lat_warm_all(l: *mut SfLat, scene: c_int) -> c_int;
        pub fn sf_lat_wake_all(l: *mut SfLat, scene: c_int) -> c_int;
        pub fn sf_lat_explode(l: *mut SfLat, scene: c_int, x: f32, y: f32, z: f32, r: f32, p: f32) -> c_int;
        pub fn sf_lat_step(l: *mut SfLat, dt: f32) -> c_int;
        pub fn sf_lat_stats(l: *mut SfLat, scene: c_int, out16: *mut f32) -> c_int;
        pub fn sf_lat_counts(l: *mut SfLat, scene: c_int, out3: *mut i32) -> c_int;
        pub fn sf_lat_bodies(l: *mut SfLat, scene: c_int, out: *mut f32) -> c_int;
        pub fn sf_lat_bricks(l: *mut SfLat, scene: c_int, out: *mut f32) -> c_int;
        pub fn sf_lat_structs(l: *mut SfLat, scene: c_int, out: *mut f32) -> c_int;
        pub fn sf_lat_matstats(l: *mut SfLat, scene: c_int, out50: *mut f32) -> c_int;
        pub fn sf_lat_set_field(l: *mut SfLat, scene: c_int, structure: c_int, phi: *const f32) -> c_int;
        pub fn sf_lat_failures(l: *mut SfLat, scene: c_int, out4: *mut u32) -> c_int;
        pub fn sf_lat_bytes_resident(l: *mut SfLat, scene: c_int) -> i64;
        pub fn sf_lat_phase_ms(l: *mut SfLat, out8: *mut f64) -> c_int;
        pub fn sf_lat_subphase_ms(l: *mut SfLat, out16: *mut f64) -> c_int;
        pub fn sf_lat_set_fused(l: *mut SfLat, max_contacts: c_int) -> c_int;
        pub fn sf_lat_set_host_threads(l: *mut SfLat, n: c_int) -> c_int;
        pub fn sf_lat_set_conn_device(l: *mut SfLat, on: c_int) -> c_int;
        pub fn sf_lat_set_sor_lds(l: *mut SfLat, max_bricks: c_int) -> c_int;

        // load query (src/lib.rs sor_struct / stress_brick)
        pub fn sf_load_create() -> *mut SfLoad;
        pub fn sf_load_destroy(l: *mut SfLoad);
        pub fn sf_load_add(l: *mut SfLoad, n: c_int, mat: *const u8, xyz: *const i32, infill: *const u8,
                           ext: *const f32) -> c_int;
        pub fn sf_load_solve(l: *mut SfLoad, omega: f32, rel_tol: f64, max_sweeps: c_int) -> c_int;
        pub fn sf_load_solve_begin(l: *mut SfLoad, omega: f32, rel_tol: f64, max_sweeps: c_int) -> c_int;
        pub fn sf_load_solve_ready(l: *mut SfLoad) -> c_int;
        pub fn sf_load_solve_end(l: *mut SfLoad) -> c_int;
        pub fn sf_load_info(l: *mut SfLoad, sid: c_int, sweeps: *mut i32, resid: *mut f64, grounded: *mut i32) -> c_int;
        pub fn sf_load_field(l: *mut SfLoad, sid: c_int, phi: *mut f64) -> c_int;
        pub fn sf_load_verdicts(l: *mut SfLoad, sid: c_int, ratio4: *mut f64, mode4: *mut u8, fail4: *mut u8) -> c_int;
        pub fn sf_load_solve_static(n: c_int, mat: *const u8, xyz: *const i32, infill: *const u8, ext: *const f32,
                                    omega: f32, rel_tol: f64, max_sweeps: c_int, stream: *mut c_void, phi: *mut f64,
                                    ratio4: *mut f64, mode4: *mut u8, fail4: *mut u8, sweeps: *mut i32,
                                    resid: *mut f64) -> c_int;

        // served what-if stream
        pub fn sf_serve_create(slots: c_int, pool_bricks: c_int, slot_bodies: c_int, frames: c_int, hbm_budget: i64,
                               max_solving: c_int) -> *mut SfServe;
        pub fn sf_serve_destroy(s: *mut SfServe);
        pub fn sf_serve_submit(s: *mut SfServe, id: u64, n: c_int, xyz: *const i32, mat: *const u8, infill: *const u8,
                               ncharge: c_int, charges5: *const f32, seed: u32) -> c_int;
        pub fn sf_serve_tick(s: *mut SfServe, dt: f32) -> c_int;
        pub fn sf_serve_poll(s: *mut SfServe, out: *mut SfAnswer, cap: c_int) -> c_int;
        pub fn sf_serve_status(s: *mut SfServe, out8: *mut i64) -> c_int;
        pub fn sf_serve_timing(s: *mut SfServe, out4: *mut f64) -> c_int;

        // the legacy engine's own exports on the default world
        pub fn sf_legacy_init() -> u32;
        pub fn sf_legacy_step(dt: f32);
        pub fn sf_legacy_explode(x: f32, y: f32, z: f32, r: f32, p: f32);
        pub fn sf_legacy_throw_charge(ox: f32, oy: f32, oz: f32, vx: f32, vy: f32, vz: f32) -> u32;
        pub fn sf_legacy_detonate() -> u32;
        pub fn sf_legacy_n_structs() -> u32;
        pub fn sf_legacy_ptr_structs() -> *const f32;
        pub fn sf_legacy_ptr_stats() -> *const f32;
        pub fn sf_legacy_clear_events();
        pub fn sf_legacy_n_events() -> u32;
        pub fn sf_legacy_ptr_events() -> *const f32;
        pub fn sf_legacy_dbg_hist(s: u32) -> *const f32;
        pub fn sf_legacy_dbg_bodies() -> *const f32;
        pub fn sf_legacy_dbg_matstats() -> *const f32;
        pub fn sf_legacy_dbg_speed() -> *const f32;
        pub fn sf_legacy_dbg_bigbody() -> *const f32;
        pub fn sf_legacy_dbg_top(s: u32) -> *const f32;
        pub fn sf_legacy_dbg_wake_all();
        pub fn sf_legacy_world(scene: *mut c_int) -> *mut SfLat;
        // structural analysis (engine/src/structure.rs over engine/src/brick.rs)
        pub fn sf_bricks_create() -> *mut SfBricks;
        pub fn sf_bricks_destroy(b: *mut SfBricks);
        pub fn sf_bricks_len(b: *const SfBricks) -> c_int;
        pub fn sf_bricks_add(b: *mut SfBricks, min3: *const i32, size3: *const i32, mat: c_int, paint: c_int, structure: c_int) -> c_int;
        pub fn sf_bricks_set_flags(b: *mut SfBricks, i: c_int, flags: c_int) -> c_int;
        pub fn sf_bricks_flags(b: *const SfBricks, i: c_int) -> c_int;
        pub fn sf_bricks_mass(b: *const SfBricks, i: c_int) -> f32;
        pub fn sf_bricks_remove(b: *mut SfBricks, i: c_int) -> c_int;
        pub fn sf_bricks_build_joints(b: *mut SfBricks) -> c_int;
        pub fn sf_bricks_joints(b: *const SfBricks, out5: *mut i32, cap: c_int) -> c_int;
        pub fn sf_bricks_set_joint_damage(b: *mut SfBricks, j: c_int, dmg: c_int) -> c_int;
        // brick.rs Bricks::save / Bricks::load (:116-145)
        pub fn sf_bricks_save(b: *const SfBricks, out: *mut u8, cap: i64) -> i64;
        pub fn sf_bricks_load(src: *const u8, n: i64) -> *mut SfBricks;
        pub fn sf_bricks_set_cell(b: *mut SfBricks, i: c_int, body: c_int, damage: c_int, hull: c_int) -> c_int;
        pub fn sf_bricks_cell(b: *const SfBricks, i: c_int, out3: *mut i32) -> c_int;
        pub fn sf_bricks_box(b: *const SfBricks, i: c_int, out6: *mut i32) -> c_int;
        // convex chunks (engine/src/hull.rs; gpu/gpu_world.js hullLibrary, buildBodies, GpuWorld.hullShapes)
        pub fn sf_build_body(n: c_int, shapes_in: *const f32, fixed: c_int, v: *const f32, w: *const f32, body: *mut f32, shapes: *mut f32) -> c_int;
        pub fn sf_hull_count() -> c_int;
        pub fn sf_hull_pattern_count() -> c_int;
        pub fn sf_hull_pattern(p: c_int, ids: *mut i32) -> c_int;
        pub fn sf_hull_cell(h: c_int, out26: *mut f64) -> c_int;
        pub fn sf_hull_vertices(h: c_int, cube: c_int, out: *mut f64) -> c_int;
        pub fn sf_hull_face(h: c_int, f: c_int, cube: c_int, xyz: *mut f64, plane4: *mut f64, other: *mut i32) -> c_int;
        pub fn sf_hull_edges(h: c_int, pairs: *mut i32, dirs: *mut f32) -> c_int;
        pub fn sf_hull_shapes(p: *const f32, q: *const f32, h: *const f32, pattern: c_int, mat: c_int, out: *mut f32) -> c_int;
        pub fn sf_warm_create() -> *mut SfWarm;
        pub fn sf_warm_destroy(w: *mut SfWarm);
        pub fn sf_warm_size(w: *const SfWarm) -> c_int;
        pub fn sf_net_create() -> *mut SfNet;
        pub fn sf_net_destroy(n: *mut SfNet);
        pub fn sf_net_add(n: *mut SfNet, b: *const SfBricks, count: c_int, set: *const u32, load: *const f32, sink: *const u8,
                          dir3: *const f32, npoint: c_int, point4: *const f32, warm: *mut SfWarm) -> c_int;
        pub fn sf_net_solve(n: *mut SfNet) -> c_int;
        pub fn sf_net_clear(n: *mut SfNet) -> c_int;
        pub fn sf_net_count(n: *const SfNet) -> c_int;
        pub fn sf_net_set_host_threads(n: *mut SfNet, threads: c_int) -> c_int;
        pub fn sf_net_set_route(n: *mut SfNet, min_cells: c_int, max_cases: c_int, block_its: c_int) -> c_int;
        pub fn sf_net_result(n: *const SfNet, c: c_int, out12: *mut f64) -> c_int;
        pub fn sf_net_levels(n: *const SfNet, c: c_int, out8: *mut i32) -> c_int;
        pub fn sf_net_joints(n: *const SfNet, c: c_int, jid: *mut i32, ratio: *mut f32, cap: c_int) -> c_int;
        pub fn sf_net_inner(n: *const SfNet, c: c_int, bat3: *mut i32, ratio: *mut f32, cap: c_int) -> c_int;
        pub fn sf_net_cells(n: *const SfNet, c: c_int, cell7: *mut i32, x6: *mut f64, cap: c_int) -> c_int;
        pub fn sf_net_ms(n: *const SfNet, out4: *mut f64) -> c_int;
        pub fn sf_net_host_ms(n: *const SfNet, out4: *mut f64) -> c_int;
    }
}

/// A failed call: the library's code and message.
#[derive(Debug, Clone)]
pub struct SfError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for SfError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "libsf_hip error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for SfError {}

fn last_error(code: i32) -> SfError {
    let msg = unsafe {
        let p = ffi::sf_last_error_message();
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    };
    SfError { code, message: msg }
}

fn check(rc: c_int) -> Result<i32, SfError> {
    if rc < 0 { Err(last_error(rc)) } else { Ok(rc) }
}

/// The brick world: one or many scenes stepping together (`sf_lat_*`).
pub struct Lattice {
    h: *mut ffi::SfLat,
}

impl Lattice {
    pub fn new(brick_cap: i32, body_cap: i32, max_scenes: i32) -> Result<Self, SfError> {
        let h = unsafe { ffi::sf_lat_create(brick_cap, body_cap, max_scenes, 16, 8) };
        if h.is_null() { Err(last_error(unsafe { ffi::sf_last_error() })) } else { Ok(Lattice { h }) }
    }
    pub fn add_scene(&mut self, brick_cap: i32, body_cap: i32, max_structs: i32, seed: u32) -> Result<i32, SfError> {
        check(unsafe { ffi::sf_lat_scene_add(self.h, brick_cap, body_cap, max_structs, seed) })
    }
    /// a standing structure: lattice cells (3 per brick), materials 1..5, infill flags
    pub fn add_structure(&mut self, scene: i32, xyz: &[i32], mat: &[u8], infill: &[u8]) -> Result<i32, SfError> {
        assert!(xyz.len() == 3 * mat.len() && infill.len() == mat.len());
        check(unsafe { ffi::sf_lat_add_structure(self.h, scene, mat.len() as c_int, xyz.as_ptr(), mat.as_ptr(), infill.as_ptr()) })
    }
    pub fn finalize(&mut self, scene: i32) -> Result<(), SfError> {
        check(unsafe { ffi::sf_lat_finalize(self.h, scene) }).map(|_| ())
    }
    pub fn warm_all(&mut self, scene: i32) -> Result<(), SfError> {
        check(unsafe { ffi::sf_lat_warm_all(self.h, scene) }).map(|_| ())
    }
    pub fn explode(&mut self, scene: i32, p: [f32; 3], r: f32, power: f32) -> Result<(), SfError> {
        check(unsafe { ffi::sf_lat_explode(self.h, scene, p[0], p[1], p[2], r, power) }).map(|_| ())
    }
    /// one frame of every scene (src/lib.rs step)
    pub fn step(&mut self, dt: f32) -> Result<(), SfError> {
        check(unsafe { ffi::sf_lat_step(self.h, dt) }).map(|_| ())
    }
    /// the ptr_stats layout (src/lib.rs:3718-3720)
    pub fn stats(&self, scene: i32) -> Result<[f32; 16], SfError> {
        let mut out = [0.0f32; 16];
        check(unsafe { ffi::sf_lat_stats(self.h, scene, out.as_mut_ptr()) })?;
        Ok(out)
    }
    pub fn failures(&self, scene: i32) -> Result<[u32; 4], SfError> {
        let mut out = [0u32; 4];
        check(unsafe { ffi::sf_lat_failures(self.h, scene, out.as_mut_ptr()) })?;
        Ok(out)
    }
    pub fn bytes_resident(&self, scene: i32) -> Result<i64, SfError> {
        let b = unsafe { ffi::sf_lat_bytes_resident(self.h, scene) };
        if b < 0 { Err(last_error(b as i32)) } else { Ok(b) }
    }
}

impl Drop for Lattice {
    fn drop(&mut self) {
        unsafe { ffi::sf_lat_destroy(self.h) }
    }
}

/// One structure's load-path verdicts: per brick four bonds (+x, +y, +z, footing).
pub struct Verdicts {
    pub ratio: Vec<f64>,
    pub mode: Vec<u8>,
    pub fail: Vec<u8>,
}

/// A batch of structures solved together (`sf_load_*`).
pub struct LoadBatch {
    h: *mut ffi::SfLoad,
    sizes: Vec<usize>,
}

impl LoadBatch {
    pub fn new() -> Self {
        LoadBatch { h: unsafe { ffi::sf_load_create() }, sizes: Vec::new() }
    }
    pub fn add(&mut self, xyz: &[i32], mat: &[u8], infill: &[u8]) -> Result<i32, SfError> {
        assert!(xyz.len() == 3 * mat.len() && infill.len() == mat.len());
        let sid = check(unsafe { ffi::sf_load_add(self.h, mat.len() as c_int, mat.as_ptr(), xyz.as_ptr(), infill.as_ptr(), core::ptr::null()) })?;
        self.sizes.push(mat.len());
        Ok(sid)
    }
    /// to max|b - A phi| / max|b| <= rel_tol in FP64; returns the structures that converged
    pub fn solve(&mut self, rel_tol: f64) -> Result<i32, SfError> {
        check(unsafe { ffi::sf_load_solve(self.h, 0.0, rel_tol, 200_000) })
    }
    pub fn field(&self, sid: i32) -> Result<Vec<f64>, SfError> {
        let mut phi = vec![0.0f64; self.sizes[sid as usize]];
        check(unsafe { ffi::sf_load_field(self.h, sid, phi.as_mut_ptr()) })?;
        Ok(phi)
    }
    pub fn verdicts(&self, sid: i32) -> Result<Verdicts, SfError> {
        let n = 4 * self.sizes[sid as usize];
        let mut v = Verdicts { ratio: vec![0.0; n], mode: vec![0; n], fail: vec![0; n] };
        check(unsafe { ffi::sf_load_verdicts(self.h, sid, v.ratio.as_mut_ptr(), v.mode.as_mut_ptr(), v.fail.as_mut_ptr()) })?;
        Ok(v)
    }
}

impl Default for LoadBatch {
    fn default() -> Self {
        Self::new()
    }
}

impl Drop for LoadBatch {
    fn drop(&mut self) {
        unsafe { ffi::sf_load_destroy(self.h) }
    }
}

/// The continuous-batching what-if server (`sf_serve_*`).
pub struct Server {
    h: *mut ffi::SfServe,
}

impl Server {
    pub fn new(slots: i32, pool_bricks: i32, slot_bodies: i32, frames: i32, hbm_budget: i64) -> Result<Self, SfError> {
        let h = unsafe { ffi::sf_serve_create(slots, pool_bricks, slot_bodies, frames, hbm_budget, 3) };
        if h.is_null() { Err(last_error(unsafe { ffi::sf_last_error() })) } else { Ok(Server { h }) }
    }
    /// charges: (x, y, z, radius, power) each
    pub fn submit(&mut self, id: u64, xyz: &[i32], mat: &[u8], infill: &[u8], charges: &[[f32; 5]], seed: u32) -> Result<(), SfError> {
        assert!(xyz.len() == 3 * mat.len() && infill.len() == mat.len());
        check(unsafe {
            ffi::sf_serve_submit(self.h, id, mat.len() as c_int, xyz.as_ptr(), mat.as_ptr(), infill.as_ptr(),
                                 charges.len() as c_int, charges.as_ptr() as *const f32, seed)
        })
        .map(|_| ())
    }
    /// one admission, activation, batched frame and retirement; returns the live scenes
    pub fn tick(&mut self, dt: f32) -> Result<i32, SfError> {
        check(unsafe { ffi::sf_serve_tick(self.h, dt) })
    }
    pub fn poll(&mut self) -> Result<Vec<ffi::SfAnswer>, SfError> {
        let mut buf = vec![ffi::SfAnswer::default(); 64];
        let mut out = Vec::new();
        loop {
            let n = check(unsafe { ffi::sf_serve_poll(self.h, buf.as_mut_ptr(), buf.len() as c_int) })? as usize;
            out.extend_from_slice(&buf[..n]);
            if n < buf.len() {
                return Ok(out);
            }
        }
    }
}

impl Drop for Server {
    fn drop(&mut self) {
        unsafe { ffi::sf_serve_destroy(self.h) }
    }
}

/// The legacy engine's exports under their reference names (src/lib.rs:3614-4053, tools/debug_exports.rs), on the
/// library's default world. Pointer results stay valid until the next call, as the wasm exports' do.
pub mod legacy {
    use super::ffi;

    pub fn init() -> u32 {
        unsafe { ffi::sf_legacy_init() }
    }
    pub fn step(dt: f32) {
        unsafe { ffi::sf_legacy_step(dt) }
    }
    pub fn explode(x: f32, y: f32, z: f32, r: f32, p: f32) {
        unsafe { ffi::sf_legacy_explode(x, y, z, r, p) }
    }
    pub fn throw_charge(ox: f32, oy: f32, oz: f32, vx: f32, vy: f32, vz: f32) -> u32 {
        unsafe { ffi::sf_legacy_throw_charge(ox, oy, oz, vx, vy, vz) }
    }
    pub fn detonate() -> u32 {
        unsafe { ffi::sf_legacy_detonate() }
    }
    pub fn n_structs() -> u32 {
        unsafe { ffi::sf_legacy_n_structs() }
    }
    pub fn ptr_structs() -> *const f32 {
        unsafe { ffi::sf_legacy_ptr_structs() }
    }
    pub fn ptr_stats() -> *const f32 {
        unsafe { ffi::sf_legacy_ptr_stats() }
    }
    pub fn clear_events() {
        unsafe { ffi::sf_legacy_clear_events() }
    }
    pub fn n_events() -> u32 {
        unsafe { ffi::sf_legacy_n_events() }
    }
    pub fn ptr_events() -> *const f32 {
        unsafe { ffi::sf_legacy_ptr_events() }
    }
    pub fn dbg_hist(s: u32) -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_hist(s) }
    }
    pub fn dbg_bodies() -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_bodies() }
    }
    pub fn dbg_matstats() -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_matstats() }
    }
    pub fn dbg_speed() -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_speed() }
    }
    pub fn dbg_bigbody() -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_bigbody() }
    }
    pub fn dbg_top(s: u32) -> *const f32 {
        unsafe { ffi::sf_legacy_dbg_top(s) }
    }
    pub fn dbg_wake_all() {
        unsafe { ffi::sf_legacy_dbg_wake_all() }
    }
    /// ptr_stats as a slice (16 values)
    pub fn stats() -> [f32; 16] {
        let mut out = [0.0f32; 16];
        let p = ptr_stats();
        if !p.is_null() {
            unsafe { out.copy_from_slice(core::slice::from_raw_parts(p, 16)) }
        }
        out
    }
}
