"""The STD-deep (task_id 4) quantum-jump tree, derived from the reference's node
recursion rather than from the runner, so that tests can check the runner and
the oracle against it.

init_split_structure_dimer_deep (core.cpp:9928-9981) gives each drive branch a
head node of dt = 0.5 T / deep_num_steps with two children.  init_split_branches
(qj_proc.cpp:4-28) sets a child's dt from its parent's: dt = parent_dt * 0.5, then
*0.5 again before every level, each level holding 2 steps.  one_sub_period_deep
(qj_proc.cpp:345-360) runs both children, and ex_period_obs_deep (core.cpp:114-153)
runs deep_num_steps sub-periods per drive branch.
"""


def reference_deep_tree(T, deep_num_steps, ex_deep):
    head_dt = 0.5 * T / float(deep_num_steps)
    children = []
    for _ in range(2):
        dt = head_dt * 0.5
        levels = []
        for _ in range(ex_deep):
            dt *= 0.5
            levels.append((2, dt))
        children.append(levels)
    level0_steps_per_sub = sum(levels[0][0] for levels in children)
    sub_per_period = 2 * deep_num_steps
    return {
        "children": children,
        "level_dt": [dt for _, dt in children[0]],
        "level0_steps_per_sub": level0_steps_per_sub,
        "sub_per_period": sub_per_period,
        # level-0 steps per period of the generalised tree (S/2 per half period)
        "S": level0_steps_per_sub * sub_per_period,
    }
