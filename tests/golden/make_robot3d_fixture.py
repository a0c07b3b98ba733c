"""Extracts the robot_3d_localization measurement tables from the reference's generated data file
(symforce/examples/robot_3d_localization/gen/measurements.cc) into a JSON fixture of numbers.

Only numbers are kept (no source text): body_t_landmark_measurements [5][20][3],
odometry_relative_pose_measurements [4][7] (Pose3 storage [qx qy qz qw tx ty tz]) and
landmark_positions [20][3]. Known answers pinned with this fixture come from
test/symforce_examples_robot_3d_localization_test.py:49-51 (initial error 463700.5576620833,
final error < 140, SUCCESS). Run from the repo root:  python tests/golden/make_robot3d_fixture.py
"""
import json
import re
from pathlib import Path

SRC = Path("/root/reference/symforce/examples/robot_3d_localization/gen/measurements.cc")
OUT = Path(__file__).resolve().parent / "robot_3d_localization.json"

NUM = r"-?\d+\.\d+(?:e-?\d+)?"


def section(text, name):
    start = text.index(name)
    end = text.index("};", start)
    return text[start:end]


def main():
    text = SRC.read_text()
    meas = [float(x) for x in re.findall(NUM, section(text, "body_t_landmark_measurements"))]
    odo = [float(x) for x in re.findall(NUM, section(text, "odometry_relative_pose_measurements"))]
    lms = [float(x) for x in re.findall(NUM, section(text, "landmark_positions{"))]
    assert len(meas) == 5 * 20 * 3 and len(odo) == 4 * 7 and len(lms) == 20 * 3
    data = {
        "num_poses": 5,
        "num_landmarks": 20,
        "body_t_landmark_measurements": [[meas[(i * 20 + j) * 3:(i * 20 + j) * 3 + 3]
                                          for j in range(20)] for i in range(5)],
        "odometry_relative_pose_measurements": [odo[i * 7:i * 7 + 7] for i in range(4)],
        "landmark_positions": [lms[j * 3:j * 3 + 3] for j in range(20)],
        "odometry_diagonal_sigmas": [0.05, 0.05, 0.05, 0.2, 0.2, 0.2],
        "matching_sigma": 0.1,
        "params": {"initial_lambda": 1e4, "lambda_down_factor": 0.5},
        "known_answers": {"initial_error": 463700.5576620833, "final_error_below": 140.0,
                          "status": "SUCCESS"},
    }
    OUT.write_text(json.dumps(data, indent=1))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
