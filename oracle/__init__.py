"""CPU oracle for the WarpGBM training / inference hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` legs of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline); ``__graft_entry__.build()`` compiles
its C restatement.  The product path (``warpgbm_amd``)
never imports it and fails loudly when its HIP library is missing.

``cpu_gbm`` (``cpu_gbm.c`` through ctypes, built by ``make -C oracle``) is a
multithreaded FP64 C restatement of the tree growth and the regression and
softmax boosting loops, the timed CPU baseline; tests/test_oracle.py pins it
against ``gbm_oracle`` (same trees, leaves and row -> leaf maps, scores within
1e-9).

``gbm_oracle`` is an FP64 NumPy restatement of the reference's kernel and
host semantics (each function cites the reference file:line it follows).
Bin edges and bin codes are defined in float32, bit for bit; everything
else (histograms, gains, leaf values, predictions) is carried in float64.

Pinning status of each function (what checks it besides itself):
  linspace_f32 / quantile_edges_f32 -- torch.linspace / torch.quantile (CPU,
      float32) over every num_bins 2..127 and several N, NaN and ties.
  bin_codes                         -- np.searchsorted(edges, x, 'right').
  histogram                         -- np.bincount per feature/era, and the
      invariants sum_b G = sum_rows g, parent = left + right.
  split_gains                       -- brute force: gain recomputed from row
      subsets by direct summation on tiny inputs.
  select_split                      -- brute-force enumeration of all (f, b)
      with the selection rule written as a plain Python loop; hand-built
      special cases (conflicting eras -> leaf, mcw > n -> leaf).
  grow_tree / partition / leaf      -- row->leaf assignment cross-checked by
      traversing the produced tree with predict_tree; leaf = closed form
      mean of g over h on each leaf's rows.
  flatten_tree / predict_forest     -- traversal of the nested dict vs the
      flattened array; chunk invariance.
  softmax_grad_hess                 -- closed form on 2-class inputs
      (sigmoid) and rows of the softmax summing to 1.
  era_inverse                       -- torch.unique(..., return_inverse=True)
      on the CPU (the routine the reference calls).
  fit_regression                    -- closed form of single-leaf trees
      (min_child_weight above the row count): F_t = ybar_w + (ybar -
      ybar_w) (1 - lr S_w / (S_w + lambda))^t with the unweighted base ybar.
  per_era_importance                -- brute force: each split's era-e gain
      recomputed from the node's era-e rows directly; one era equals the
      total gain.
  predict_proba_binned              -- K = 2 sigmoid closed form of the two
      class forests, rows summing to 1, class-permutation equivariance.
  fit_multi_regression / grow_tree_multi / multi_regression_targets /
  predict_multi                     -- one output with unit weights equals
      fit_regression; an all-zero-weight output leaves the trees of the
      other output's regression fit and gets zero leaves and base; NaN
      targets equal zero weights; base by a row-by-row weighted sum; output
      permutation equivariance; each leaf vector by direct sums over the
      rows reaching it; predict_multi reproduces the fit's scores.
"""
