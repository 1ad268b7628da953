function [out, flag, relres, iter, resvec] = mobilityProblem(swimmers, fluidFunc, t, conditions, ...
    kernalPar, TreePar, threads, GPU, blockSize, tol, maxIter, PreconSettings)
%MOBILITYPROBLEM Velocities and basis rates [U; Omega x b1; Omega x b2] of
%   every body (Mobility/mobilityProblem.m:24-72), with the GMRES of the
%   saddle-point system run on the GPU (unrestarted, from x0 = 0,
%   preconditioned on the right when PreconSettings is not empty).  The
%   extra outputs are gmres's flag, relres, iter and resvec, which the
%   reference drops.
for i = 0:(numel(conditions) / 9) - 1
    data = conditions(i * 9 + 1:(i + 1) * 9);
    swimmers.updateSwimmmer(i + 1, data(1:3)', data(4:9));
end
[out, flag, relres, iter, resvec] = nnkSolveMobility(swimmers, fluidFunc, t, conditions, kernalPar, ...
    TreePar, threads, GPU, blockSize, tol, maxIter, PreconSettings);
end
