"""Poses: nalgebra Isometry3<f32> as the 32-byte ncx_pose record (t, pad, q_ijkw)."""
import numpy as np


def poses_array(translations, quats_ijkw):
    """(n, 8) float32 array in ncx_pose layout from translations (n,3) and unit quaternions (n,4: i,j,k,w)."""
    t = np.asarray(translations, np.float32).reshape(-1, 3)
    q = np.asarray(quats_ijkw, np.float32).reshape(-1, 4)
    out = np.zeros((t.shape[0], 8), np.float32)
    out[:, 0:3] = t
    out[:, 4:8] = q
    return out


def identity_poses(n):
    return poses_array(np.zeros((n, 3)), np.tile([0, 0, 0, 1], (n, 1)))


def to_oracle_poses(poses8):
    """(n, 8) ncx_pose layout -> (n, 7) (tx, ty, tz, qi, qj, qk, qw)."""
    p = np.asarray(poses8).reshape(-1, 8)
    return np.concatenate([p[:, 0:3], p[:, 4:8]], 1)
