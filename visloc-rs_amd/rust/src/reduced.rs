//! `ReducedSolveAccelerator` (pipelines/slam/src/ba_accel.rs:134-157) on
//! `vba_reduced_solve` and `vba_reduced_eliminate`, for the COLMAP-port rig
//! solver's iterative reduced-camera branch (rig_ba_solver.rs:2932-3101).
//! nalgebra's `Matrix6` and `DMatrix` are column-major, the layout the C ABI
//! takes, so blocks and inverses pass through as flat slices.
use nalgebra::{Matrix6, Vector6};
use visloc_slam::{EliminatedBlocks, PointElimination, ReducedCameraSystem, ReducedSolveAccelerator};

use crate::{ffi, VbaAccelerator};

impl ReducedSolveAccelerator for VbaAccelerator {
    fn solve(&self, sys: &ReducedCameraSystem<'_>) -> Option<(Vec<Vector6<f64>>, usize)> {
        let n = sys.diag.len();
        // an empty offdiag: the blocks the last `eliminate` of this structure
        // left on the device
        if !sys.offdiag.is_empty() && sys.offdiag.len() != 36 * sys.edges.len() {
            return None;
        }
        let offdiag = if sys.offdiag.is_empty() { std::ptr::null() } else { sys.offdiag.as_ptr() };
        let diag: Vec<f64> = sys.diag.iter().flat_map(|d| d.as_slice().iter().copied()).collect();
        let edges: Vec<u32> = sys.edges.iter().flat_map(|&(a, b)| [a, b]).collect();
        let rhs: Vec<f64> = sys.rhs.iter().flat_map(|v| v.iter().copied()).collect();
        // runs of equal length from node 0, as cluster_preconditioner builds them
        let (cluster_size, cluster_inverses): (u32, Vec<f64>) = match sys.clusters {
            Some(c) => {
                let k = c.first()?.1.nrows() / 6;
                if k == 0 || c.iter().enumerate().any(|(i, (s, m))| *s != i * k || m.nrows() != 6 * k) {
                    return None;
                }
                (k as u32, c.iter().flat_map(|(_, m)| m.as_slice().iter().copied()).collect())
            }
            None => (0, Vec::new()),
        };
        let (coarse_group, coarse_nodes, coarse_inverse) = match sys.coarse {
            Some((g, nodes, inv)) => (u32::try_from(g).ok()?, u32::try_from(nodes).ok()?, inv.as_slice().as_ptr()),
            None => (0, 0, std::ptr::null()),
        };
        let system = ffi::vba_reduced_system {
            structure_id: sys.structure_id,
            num_nodes: u32::try_from(n).ok()?,
            num_edges: u32::try_from(sys.edges.len()).ok()?,
            diag: diag.as_ptr(),
            edges: edges.as_ptr(),
            offdiag,
            rhs: rhs.as_ptr(),
            cluster_size,
            coarse_group,
            cluster_inverses: if cluster_inverses.is_empty() { std::ptr::null() } else { cluster_inverses.as_ptr() },
            coarse_nodes,
            max_iterations: u32::try_from(sys.max_iterations).unwrap_or(u32::MAX),
            coarse_inverse,
            tol: sys.tol,
        };
        let mut step = vec![0.0f64; 6 * n];
        let mut iterations = 0u32;
        let st = {
            let g = self.lock()?;
            unsafe { ffi::vba_reduced_solve(g.0, &system, step.as_mut_ptr(), &mut iterations) }
        };
        (st == ffi::VBA_OK)
            .then(|| (step.chunks_exact(6).map(Vector6::from_column_slice).collect(), iterations as usize))
    }

    fn eliminate(&self, e: &PointElimination<'_>) -> Option<EliminatedBlocks> {
        let num_nodes = e.row_off.len().checked_sub(1)?;
        let num_points = e.point_off.len().checked_sub(1)?;
        let edges: Vec<u32> = e.edges.iter().flat_map(|&(a, b)| [a, b]).collect();
        let row_off: Vec<u64> = e.row_off.iter().map(|&v| v as u64).collect();
        let slot_off: Vec<u64> = e.slot_off.iter().map(|&v| v as u64).collect();
        let slots: Vec<u32> = e.slots.iter().flat_map(|s| *s).collect();
        let point_off: Vec<u64> = e.point_off.iter().map(|&v| v as u64).collect();
        let hcp: Vec<f64> = e.hcp.iter().flat_map(|m| m.as_slice().iter().copied()).collect();
        let point_inverse: Vec<f64> = e
            .point_inverse
            .iter()
            .flat_map(|m| match m {
                Some(m) => {
                    let mut a = [0.0f64; 9];
                    a.copy_from_slice(m.as_slice());
                    a
                }
                None => [0.0f64; 9],
            })
            .collect();
        let has_inverse: Vec<u8> = e.point_inverse.iter().map(|m| m.is_some() as u8).collect();
        let scale: Vec<f64> = e.scale.iter().flat_map(|v| v.iter().copied()).collect();
        let extra_edge: Vec<u32> = e.extra.iter().map(|(k, _)| *k).collect();
        let extra_blocks: Vec<f64> = e.extra.iter().flat_map(|(_, m)| m.as_slice().iter().copied()).collect();
        let (coarse_group, coarse_nodes) = match e.coarse {
            Some((g, nodes)) => (u32::try_from(g).ok()?, u32::try_from(nodes).ok()?),
            None => (0, 0),
        };
        let m = if coarse_group > 0 { (coarse_nodes as usize).div_ceil(coarse_group as usize) } else { 0 };
        let view = ffi::vba_point_elimination {
            structure_id: e.structure_id,
            linearization_id: e.linearization_id,
            num_nodes: u32::try_from(num_nodes).ok()?,
            num_edges: u32::try_from(e.edges.len()).ok()?,
            edges: edges.as_ptr(),
            row_off: row_off.as_ptr(),
            slot_off: slot_off.as_ptr(),
            slots: slots.as_ptr(),
            num_points: u32::try_from(num_points).ok()?,
            num_extra: u32::try_from(extra_edge.len()).ok()?,
            point_off: point_off.as_ptr(),
            point_nodes: e.point_nodes.as_ptr(),
            hcp: hcp.as_ptr(),
            point_inverse: point_inverse.as_ptr(),
            has_inverse: has_inverse.as_ptr(),
            scale: scale.as_ptr(),
            extra_edge: extra_edge.as_ptr(),
            extra_blocks: extra_blocks.as_ptr(),
            num_run_edges: u32::try_from(e.run_edges.len()).ok()?,
            coarse_group,
            run_edges: e.run_edges.as_ptr(),
            coarse_nodes,
            reserved: 0,
        };
        let mut runs = vec![0.0f64; 36 * e.run_edges.len()];
        let mut coarse = vec![0.0f64; 36 * m * m];
        let st = {
            let g = self.lock()?;
            unsafe { ffi::vba_reduced_eliminate(g.0, &view, runs.as_mut_ptr(), coarse.as_mut_ptr()) }
        };
        (st == ffi::VBA_OK).then(|| EliminatedBlocks {
            run_blocks: runs.chunks_exact(36).map(Matrix6::from_column_slice).collect(),
            coarse_sums: if m > 0 { coarse.chunks_exact(36).map(Matrix6::from_column_slice).collect() } else { Vec::new() },
        })
    }
}
