# SPDX-License-Identifier: MIT
"""numpy views of the records exchanged through the C-ABI.

They restate the reference's host/device layouts byte for byte
(src/simulation/geometry_storage.hpp:18-213, src/slang/types.slang:81-103,
src/slang/geometry.slang:143-238) because the kernel seam accepts them as
the reference's dispatch wrappers do.
"""
import numpy as np

F3 = ("<f4", 3)

STATE = np.dtype([
    ("position", *F3),
    ("orientation", "<f4", 4),
    ("linear_velocity", *F3),
    ("angular_velocity", *F3),
])
assert STATE.itemsize == 52

MOTION_STATIC, MOTION_KINEMATIC, MOTION_DYNAMIC = 0, 1, 2

RIGID_INPUT = np.dtype([                      # PmRigidInput, 52 bytes
    ("body_index", "<u4"), ("force", "<f4", 3), ("torque", "<f4", 3),
    ("linear_impulse", "<f4", 3), ("angular_impulse", "<f4", 3)])

CONTACT_EVENT = np.dtype([                    # PmContactEvent, 56 bytes
    ("type", "<u4"), ("first", "<u4"), ("second", "<u4"), ("position", "<f4", 3),
    ("normal", "<f4", 3), ("penetration", "<f4"), ("normal_impulse", "<f4"),
    ("friction_impulse", "<f4", 3)])

PARAMS = np.dtype({
    "names": ["motion", "mesh_index", "mesh_generation", "inverse_mass",
              "inverse_inertia_local", "friction", "restitution",
              "linear_damping", "angular_damping", "maximum_linear_speed",
              "maximum_angular_speed", "collision_margin", "user_data"],
    "formats": ["<u4", "<u4", "<u4", "<f4", F3, "<f4", "<f4", "<f4", "<f4",
                "<f4", "<f4", "<f4", "<u8"],
    "offsets": [0, 4, 8, 12, 16, 28, 32, 36, 40, 44, 48, 52, 56],
    "itemsize": 64,
})

HINGE = np.dtype([
    ("anchor", *F3), ("axis", *F3), ("local_anchor", *F3),
    ("inverse_mass", "<f4"), ("inverse_moment", "<f4"),
    ("present", "<u4"), ("fixed", "<u4"), ("axial", "<u4"),
    ("axial_rotation", "<u4"), ("fixed_member", "<u4"),
    ("point_member", "<u4"), ("static_body", "<u4"),
])
assert HINGE.itemsize == 72

CONTACT = np.dtype([
    ("normal", *F3), ("point", *F3), ("penetration", "<f4"),
    ("body_hinge", HINGE), ("collider_hinge", HINGE),
    ("accumulated_normal_impulse", "<f4"),
    ("accumulated_friction_impulse", *F3),
    ("initial_normal_speed", "<f4"), ("warm_started", "<u4"),
    ("persistent", "<u4"), ("impact_fraction", "<f4"),
    ("reported_normal_impulse", "<f4"),
    ("reported_friction_impulse", *F3),
    ("avbd_dual", "<f4", (3, 2)), ("avbd_error", *F3),
    ("avbd_anchor_a", *F3), ("avbd_anchor_b", *F3), ("avbd_cached", "<u4"),
])
assert CONTACT.itemsize == 284

MANIFOLD = np.dtype([
    ("contacts", CONTACT, 8), ("count", "<u4"),
    ("initial_relative_position", *F3), ("face_patch", "<u4"),
    ("response_slot", "<u4"), ("avbd_next_a", "<u4"),
    ("avbd_next_b", "<u4"),
])
assert MANIFOLD.itemsize == 2304

AVBD_BODY = np.dtype([
    ("target", STATE), ("orientation_delta", "<f4", 4),
    ("displacement", "<f4", 6), ("inertia_h", "<f4", 21),
    ("inertia_g", "<f4", 6), ("contact_head", "<u4"), ("joint_head", "<u4"),
    ("color", "<u4"), ("parent", "<u4"), ("impact", "<u4"),
    ("previous_velocity", *F3), ("history_valid", "<u4"),
])
assert AVBD_BODY.itemsize == 236

SCHEDULE = np.dtype([
    ("counts", "<u4", 29), ("offsets", "<u4", 29), ("budget", "<u4"),
    ("remaining", "<u4"), ("island_count", "<u4"),
    ("early_exit_count", "<u4"), ("maximum_passes", "<u4"),
    ("blocks", "<u4"), ("candidates", "<u4"), ("contacts", "<u4"),
])
assert SCHEDULE.itemsize == 264

INVALID = 0xFFFFFFFF

# Joints (geometry.slang:24-141).
JOINT_FIXED, JOINT_POINT, JOINT_HINGE, JOINT_SLIDER, JOINT_PISTON = 0, 1, 2, 3, 4
JOINT_GENERIC, JOINT_GENERIC_SPRING, JOINT_MOTOR = 5, 6, 7

AVBD_ROW = np.dtype([                          # RigidAvbdRow, 88 bytes
    ("a", "<f4", 6), ("b", "<f4", 6), ("lambda", "<f4"), ("penalty", "<f4"),
    ("error", "<f4"), ("velocity", "<f4"), ("lower", "<f4"), ("upper", "<f4"),
    ("stiffness", "<f4"), ("damping", "<f4"), ("beta", "<f4"), ("reference_force", "<f4")])

_LIMITS = [("axes", "<u4"), ("lower", "<f4", 3), ("upper", "<f4", 3)]
_SPRINGS = [("axes", "<u4"), ("stiffness", "<f4", 3), ("damping", "<f4", 3)]

JOINT_OPTIONS = np.dtype([                     # RigidConstraintOptions, 228 bytes
    ("type", "<u4"), ("body_a", "<u4", 2), ("body_b", "<u4", 2),
    ("local_anchor_a", "<f4", 3), ("local_anchor_b", "<f4", 3),
    ("local_orientation_a", "<f4", 4), ("local_orientation_b", "<f4", 4),
    ("linear_limits", _LIMITS), ("angular_limits", _LIMITS),
    ("linear_springs", _SPRINGS), ("angular_springs", _SPRINGS),
    ("motor", [("linear_enabled", "<u4"), ("angular_enabled", "<u4"),
               ("linear_target_velocity", "<f4"), ("linear_maximum_impulse", "<f4"),
               ("angular_target_velocity", "<f4"), ("angular_maximum_impulse", "<f4")]),
    ("enabled", "<u4"), ("disable_collisions", "<u4"),
    ("breaking_impulse_threshold", "<f4"), ("solver_iterations", "<u4")])

JOINT = np.dtype([                             # RigidConstraintResource, 1852 bytes
    ("options", JOINT_OPTIONS), ("enabled", "<u4"), ("broken", "<u4"),
    ("applied_impulse", "<f4"), ("generation", "<u4"), ("alive", "<u4"),
    ("dense_a", "<u4"), ("dense_b", "<u4"), ("rows", AVBD_ROW, 18),
    ("avbd_count", "<u4"), ("avbd_next_a", "<u4"), ("avbd_next_b", "<u4")])

assert JOINT_OPTIONS.itemsize == 228 and JOINT.itemsize == 1852 and AVBD_ROW.itemsize == 88


def joint_options(n, kind, a, b, anchor_a=(0, 0, 0), anchor_b=(0, 0, 0), disable_collisions=1,
                  breaking=0.0, iterations=0):
    """Joint options between dense bodies a and b (identity local frames)."""
    o = np.zeros(n, dtype=JOINT_OPTIONS)
    o["type"] = kind
    o["body_a"][:, 0] = a
    o["body_a"][:, 1] = 1
    o["body_b"][:, 0] = b
    o["body_b"][:, 1] = 1
    o["local_anchor_a"] = anchor_a
    o["local_anchor_b"] = anchor_b
    o["local_orientation_a"] = (0, 0, 0, 1)
    o["local_orientation_b"] = (0, 0, 0, 1)
    o["enabled"] = 1
    o["disable_collisions"] = disable_collisions
    o["breaking_impulse_threshold"] = breaking
    o["solver_iterations"] = iterations
    return o

