classdef nnkOperators < handle
    %NNKOPERATORS The library operators of one KIFMM object, created on
    %   first use and shared by its copies (KIFMM is a value class, as in
    %   Classes/KIFMM.m, so computeVelSparse's coronaRes-4 operator is kept
    %   here rather than rebuilt per call).  Keys: coronaRes, precision.
    properties (SetAccess = private)
        handles;
    end
    methods
        function this = nnkOperators()
            this.handles = containers.Map('KeyType', 'char', 'ValueType', 'any');
        end
        function id = get(this, tree, kernelPar, res, shells, precision, svdTol)
            key = sprintf('%d/%d/%d/%g', res, shells, precision, svdTol);
            if ~isKey(this.handles, key)
                this.handles(key) = nnkHandle(nnkifmm_mex('fmm', tree.h.id, kernelPar(1:2), ...
                                                          res, shells, precision, svdTol));
            end
            h = this.handles(key);
            id = h.id;
        end
    end
end
