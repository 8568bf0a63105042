"""Input vectors for the turtle and spline parity tests: (name, definition, iterations). The first
group are the definitions of the reference's Application/tests/ProceduralMeshTests.cpp, the rest
exercise material switches, every rotation module, nested branches, stochastic growth and
curved spans (refinement of a span between knots whose tangents differ).

The reference turtle runs with TurtleParams{500, 100} (its defaults); the frames of every segment
are taken with samplesPerSpan 2, 4 and 7 (SplinePath::computeLoftFrames) and 4
(sampleUniform), tension 1.
"""
CASES = [
    ("single_f", "F\n", 0),
    ("f_f", "F F\n", 0),
    ("f_gap_f", "F f F\n", 0),
    ("pitch_branch", "F [Pitch(45) F]\n", 0),
    ("f_f_f", "F F F\n", 0),
    ("unbalanced_close", "F ] F ] F\n", 0),
    ("branch_groups", "F [ [ F F ] [ F F ] ] F\n", 0),
    ("yaw_chain", "F(800) Yaw(2) F(500) Yaw(1) F(300) Yaw(3) [Pitch(45) F F F][Pitch(-45) F F F]", 0),
    ("step_radius", "F(2, 1000, 1000)\n", 0),
    ("zero_step_radius", "F(0, 1000)\n", 0),
    ("zero_step", "F(0)\n", 0),
    ("materials", "Mat(0) = {1,0,0}\nMat(bark) = {0,1,0}\n"
     "F Mat(bark) F [ Mat(7) Roll(30) Pitch(20) F ] Mat(1.5) F Mat(-1) F Mat(3) F", 0),
    ("case_insensitive", "F yaw(30) F PITCH(20) F pich(10) F rOlL(45) F mat(2) F", 0),
    ("curved_chain", "F(100, 10) Yaw(30) F(100, 8) Pitch(25) F(100, 6) Roll(60) Yaw(-40) F(100, 4) Pitch(-70) F(50, 2)", 0),
    ("turn_in_place", "F Yaw(90) Yaw(-90) F Pitch(180) F", 0),
    ("plant", "Mat(0) = {0.4, 0.3, 0.2, 0.6}\nMat(1) = {0.2, 0.5, 0.1, 0.5}\nd = 0.72\nA(100, 10)\n"
     "A(l, w) : w > 1 : 0.7 -> F(l, w) [ Pitch(30) A(l*d, w*d) ] Roll(137.5) [ Pitch(-25) A(l*d, w*d) ]\n"
     "A(l, w) : w > 1 : 0.3 -> F(l, w) Roll(90) [ Yaw(20) A(l*d, w*d*d) ]\n"
     "A(l, w) : w <= 1 -> Mat(1) F(l, w)", 8),
    ("bush", "F\nF -> F [ Yaw(25) F ] F [ Yaw(-25) F ] Pitch(10) F", 3),
    ("editor_tree", "Init\nInit -> Pitch(10) F(3,0.5)\nF(h,r) -> [Yaw(22.5)F(h*0.7,r*0.55)][Yaw(0-22.5)F(h*0.7,r*0.55)]", 6),
]
