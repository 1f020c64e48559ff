"""The reference's ``atlas_q.tools_qih`` modules on this path: the second
two-site update ``tn_core.mps_apply_2q`` with its AI rank policies, the
reference's learned rank predictors (state-dict compatible, inference on the
library's dense-stack kernel), the SVD-event log and the dataset builder."""
