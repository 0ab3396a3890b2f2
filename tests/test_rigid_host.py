"""Host side of the rigid solver on the CPU: the joint tables restate NativeJointState.prepare
(Sources/BubblePhysicsNativeMetal/NativeJointState.swift:73-141) and the record staging restates
NativeRigidPolygonState.prepare (NativeRigidPolygonState.swift:193-301). No device call."""
import numpy as np

from bubblephysics_amd import abi, rigid

J = rigid.joint_desc


def test_joint_tables_sort_by_id_move_the_world_to_end_zero_and_list_partners():
    ids = np.array([5, 7, 9], np.uint64)
    joints = [J(30, 9, None, kind=abi.BP_JOINT_HINGE, anchor_a=(1, 2), anchor_b=(10, 20), limits=(-1, 1), motor=(2, 50)),
              J(10, 5, 7, length=4.0, softness=0.5, damping=3.0),
              J(20, 7, 9, kind=abi.BP_JOINT_WELD, reference_angle=0.25, collide_connected=True),
              J(40, 5, 99, length=1.0),                 # unknown body: skipped
              J(15, 7, 5, length=2.0)]                  # second joint between 5 and 7: partners stay distinct
    records, refs, ranges, partners = rigid.joint_tables(joints, ids)
    assert [int(r["id"]) for r in records] == [10, 15, 20, 30]
    r10, r15, r20, r30 = records
    assert tuple(r10["ends"]) == (0, 1, abi.BP_JOINT_DISTANCE, 0) and tuple(r10["parameters"][:3]) == (4.0, 0.5, 3.0)
    assert tuple(r20["ends"]) == (1, 2, abi.BP_JOINT_WELD, 0) and r20["parameters"][3] == np.float32(0.25)
    # A joint to the world: the world is end 0 with anchor b, the body end 1 with anchor a.
    assert tuple(r30["ends"]) == (abi.BP_JOINT_WORLD_END, 2, abi.BP_JOINT_HINGE, 3)
    assert tuple(r30["anchors"]) == (10, 20, 1, 2) and tuple(r30["limits"]) == (-1, 1, 2, 50)
    ranges = ranges.reshape(3, 4)
    refs = refs.reshape(-1, 2)
    lists = [[tuple(refs[ranges[i][0] + k]) for k in range(ranges[i][1])] for i in range(3)]
    assert lists == [[(0, 0), (1, 1)], [(0, 1), (1, 0), (2, 0)], [(2, 1), (3, 1)]]
    plist = [sorted(int(p) for p in partners[ranges[i][2]:ranges[i][2] + ranges[i][3]]) for i in range(3)]
    assert plist == [[7], [5], []], "collide_connected joints add no partners, repeated pairs appear once"


def test_records_follow_the_centroid_and_the_kinematic_rule():
    config = rigid.rigid_config(dt=0.5)
    tri = rigid.RigidObject(id=3, shape="polygon", vertices=[(0, 0), (3, 0), (0, 3)], position=(10, 20), rotation=0.0,
                            velocity=(1, 2), angular_velocity=0.5, role="kinematic")
    scene = rigid.records_for([tri], config)
    b = scene.bodies[0]
    # The centroid of the triangle is (1, 1): the record's pose is the centre of mass.
    assert np.allclose(b["pose"][:2], (11, 21)) and np.allclose(b["previous"][2:], (1, 1))
    # A kinematic body's centre-of-mass velocity adds the centroid's turn over one step.
    c, s = np.cos(0.25), np.sin(0.25)
    turn = ((1 * c - 1 * s) - 1, (1 * s + 1 * c) - 1)
    assert np.allclose(b["motion"][:2], (1 + turn[0] / 0.5, 2 + turn[1] / 0.5), atol=1e-5)
    assert b["physical"][0] == 0 and b["physical"][1] == 0 and b["filter"][2] == 1 and b["motion"][3] == 0
    pos, rot, vel, w = rigid.public_state(scene.bodies)
    assert np.allclose(pos[0], (10, 20), atol=1e-5)


def test_static_polygons_become_owned_one_sided_segments():
    config = rigid.rigid_config()
    scene = rigid.records_for([rigid.box(4, (50, 60), (10, 5), role="static", category=2, mask=6),
                               rigid.circle(5, (0, 0), 3, role="static")], config)
    segs = scene.segments
    assert len(segs) == 4 and np.all(segs["owner_id"] == 4) and np.all(segs["flags"] & abi.BP_SEGMENT_HAS_OWNER)
    assert np.all(segs["category"] == 2) and np.all(segs["mask"] == 6)
    assert scene.bodies["flags"][:, 0].tolist() == [1, 1]
