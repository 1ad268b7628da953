function [out, flag, relres, iter, resvec] = mobilityProblemGravity(swimmers, fluidFunc, t, conditions, ...
    kernalPar, TreePar, threads, GPU, blockSize, tol, maxIter, PreconSettings, gravPar)
%MOBILITYPROBLEMGRAVITY As mobilityProblem with every body carrying the
%   force G = gravPar(1:3) and the moment -(4/3) pi gravPar(4) (b3 x G)
%   (Mobility/mobilityProblemGravity.m:25-74, Classes/Swimmers.m:372-385).
for i = 0:(numel(conditions) / 9) - 1
    data = conditions(i * 9 + 1:(i + 1) * 9);
    swimmers.updateSwimmmer(i + 1, data(1:3)', data(4:9));
    swimmers.updateSwimmmerRHS(i + 1, gravPar(1:3), gravPar(4));
end
[out, flag, relres, iter, resvec] = nnkSolveMobility(swimmers, fluidFunc, t, conditions, kernalPar, ...
    TreePar, threads, GPU, blockSize, tol, maxIter, PreconSettings);
end
