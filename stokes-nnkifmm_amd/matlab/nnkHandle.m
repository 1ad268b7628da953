classdef nnkHandle < handle
    %NNKHANDLE One live libnnkifmm object (octree, KIFMM operator or
    %   mobility operator) held by its registry id.  Deleting the last
    %   MATLAB reference releases it (nnkifmm_mex('free', id)); the library
    %   keeps an object alive while operators built on it are alive.
    properties (SetAccess = private)
        id = uint64(0);
    end
    methods
        function this = nnkHandle(id)
            this.id = id;
        end
        function delete(this)
            if this.id ~= 0
                nnkifmm_mex('free', this.id);
                this.id = uint64(0);
            end
        end
    end
end
